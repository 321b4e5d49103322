! CPU ORACLE -- test infrastructure only: the harness that runs the reference's own ILMC
! loop (oracle/_ref/libilmc_ref.so; built only where the reference tree is present).
!
! The reference's sequential sweep loop, ilmc_lam_loop.inc:16-139, is a procedure contained
! in ILMC_LAM that works on its host's variables.  This subroutine declares those variables
! under the names the loop uses, fills the per-cell sweep lists (type a_ilmc of the
! reference's ilmc_lam_array.F90) the way ILMC_LAM.F90:144-214 does, and includes the loop
! from the reference tree at build time.  The min/max reset of ILMC_LAM.F90:249-273 is not
! part of it.
!
! dom = minx maxx miny maxy nk i0 in j0 jn k0 sweep_max min_max autobar (min_max unused);
! moved = the number of sweeps that moved mass (sum of reset(:,5)).
subroutine ilmc_ref_loop(f, mn, mx, air, dom, moved) bind(C, name="ilmc_ref_loop")
   use, intrinsic :: iso_c_binding
   use ilmc_lam_array
   implicit none
   type(c_ptr), value :: f, mn, mx, air
   integer(c_int), intent(in) :: dom(13)
   integer(c_long_long), intent(out) :: moved

   type post_arrays
      real, pointer, dimension(:,:,:) :: min, max
   end type post_arrays
   type(post_arrays) :: Adz_post(1)
   real, pointer, dimension(:,:,:) :: F_ilmc, air_mass_m, p3
   integer :: F_ns, Adz_k0t, l_nk, Adz_ILMC_sweep_max, il, ir, jl, jr, minx, maxx, miny, maxy
   integer :: i, j, k, sweep, i_rd, j_rd, k_rd, ii, ix, jx, kx, kx_m, kx_p, w1, w2, nc
   integer, allocatable :: reset(:,:)
   real :: sweep_max(400), sweep_min(400), m_use_max, m_use_min, m_sweep_max, m_sweep_min, &
           o_shoot, u_shoot, ratio_max, ratio_min
   logical :: autobar_L, inner_L

   minx = dom(1); maxx = dom(2); miny = dom(3); maxy = dom(4); l_nk = dom(5)
   il = dom(6); ir = dom(7); jl = dom(8); jr = dom(9); Adz_k0t = dom(10)
   Adz_ILMC_sweep_max = dom(11); autobar_L = dom(13) /= 0
   F_ns = 1

   call c_f_pointer(f, p3, [maxx-minx+1, maxy-miny+1, l_nk])
   F_ilmc(minx:, miny:, 1:) => p3
   call c_f_pointer(mn, p3, [maxx-minx+1, maxy-miny+1, l_nk])
   Adz_post(1)%min(minx:, miny:, 1:) => p3
   call c_f_pointer(mx, p3, [maxx-minx+1, maxy-miny+1, l_nk])
   Adz_post(1)%max(minx:, miny:, 1:) => p3
   call c_f_pointer(air, p3, [maxx-minx+1, maxy-miny+1, l_nk])
   air_mass_m(minx:, miny:, 1:) => p3

   ! The cells of sweep s around (i,j,k): the clipped cube of half-width s less the clipped
   ! cube of half-width s-1, visited j (outer), k, i (inner); one level with autobar.
   allocate (sweep_rd(Adz_ILMC_sweep_max, il:ir, jl:jr, l_nk))
   do k = Adz_k0t, l_nk
      do j = jl, jr
         do i = il, ir
            do sweep = 1, Adz_ILMC_sweep_max
               w1 = 2*sweep + 1
               w2 = w1 - 2
               allocate (sweep_rd(sweep,i,j,k)%i_rd(2*(w1**2 + w1*w2 + w2**2)))
               allocate (sweep_rd(sweep,i,j,k)%j_rd(2*(w1**2 + w1*w2 + w2**2)))
               allocate (sweep_rd(sweep,i,j,k)%k_rd(2*(w1**2 + w1*w2 + w2**2)))
               kx_m = k ; kx_p = k
               if (.not. autobar_L) then
                  kx_m = max(k-sweep, Adz_k0t) ; kx_p = min(k+sweep, l_nk)
               end if
               nc = 0
               do jx = max(j-sweep, jl), min(j+sweep, jr)
                  do kx = kx_m, kx_p
                     inner_L = abs(jx-j) < sweep
                     if (.not. autobar_L) inner_L = inner_L .and. abs(kx-k) < sweep
                     do ix = max(i-sweep, il), min(i+sweep, ir)
                        if (inner_L .and. abs(ix-i) < sweep) cycle
                        nc = nc + 1
                        sweep_rd(sweep,i,j,k)%i_rd(nc) = ix
                        sweep_rd(sweep,i,j,k)%j_rd(nc) = jx
                        sweep_rd(sweep,i,j,k)%k_rd(nc) = kx
                     end do
                  end do
               end do
               sweep_rd(sweep,i,j,k)%cell = nc
            end do
         end do
      end do
   end do

   allocate (reset(Adz_k0t:l_nk, 5))
   reset = 0
   call ilmc_lam_loop ()
   moved = sum(reset(:,5))

   do k = Adz_k0t, l_nk
      do j = jl, jr
         do i = il, ir
            do sweep = 1, Adz_ILMC_sweep_max
               deallocate (sweep_rd(sweep,i,j,k)%i_rd, sweep_rd(sweep,i,j,k)%j_rd, sweep_rd(sweep,i,j,k)%k_rd)
            end do
         end do
      end do
   end do
   deallocate (sweep_rd, reset)

contains

   include 'ilmc_lam_loop.inc'

end subroutine ilmc_ref_loop
