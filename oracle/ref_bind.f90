! Test-only C binding of the parts of the reference scadyn Fortran that build
! here without FFTW or HDF5 (common, integration_points,
! singularity_subtraction, sfunctions, geometry; compiled in place from the
! reference tree by oracle/build_ref.py into oracle/_ref/).  Used only by
! tests/ to check the oracle's restatement against the reference code itself.
! The two loops of rb_vv_dvdv follow integrate_V_V_G / integrate_dV_dV_G
! (src/precorrection.f90:729-806), whose module needs FFTW and cannot be built.
module ref_bind
   use iso_c_binding
   use common
   use integration_points
   use singularity_subtraction
   use sfunctions
   use geometry
   implicit none
contains

   subroutine rb_analytical(tri, nv, r, out) bind(C, name='rb_analytical')
      real(c_double), intent(in) :: tri(3, 3), nv(3), r(3)
      real(c_double), intent(out) :: out(2)
      real(dp) :: t(3, 3), n(3), p(3), res(2)
      t = tri; n = nv; p = r
      res = analytical_integration(t, n, p)
      out = res
   end subroutine

   subroutine rb_int_v_g(rf, k, pb, wb, nb, bc, out) bind(C, name='rb_int_v_g')
      integer(c_int), value :: nb
      real(c_double), value :: k
      real(c_double), intent(in) :: rf(3), pb(3, nb), wb(nb), bc(3, 4)
      complex(c_double_complex), intent(out) :: out
      real(dp), allocatable :: pbx(:, :), wbx(:), bcx(:, :)
      real(dp) :: r(3), kk
      allocate (pbx(3, nb), wbx(nb), bcx(3, 4))
      pbx = pb; wbx = wb; bcx = bc; r = rf; kk = k
      out = singularity_subtraction_int_V_G(r, kk, pbx, wbx, bcx)
   end subroutine

   subroutine rb_int_s_g(rf, tri, nv, k, pb, wb, nb, out) bind(C, name='rb_int_s_g')
      integer(c_int), value :: nb
      real(c_double), value :: k
      real(c_double), intent(in) :: rf(3), tri(3, 3), nv(3), pb(3, nb), wb(nb)
      complex(c_double_complex), intent(out) :: out
      real(dp) :: r(3), t(3, 3), n(3), kk
      real(dp), allocatable :: pbx(:, :), wbx(:)
      allocate (pbx(3, nb), wbx(nb))
      pbx = pb; wbx = wb; r = rf; t = tri; n = nv; kk = k
      out = singularity_subtraction_int_S_G(r, t, n, kk, pbx, wbx)
   end subroutine

   ! V_V scalar and the 3x3 dV_dV matrix of one (test, basis) tet pair
   subroutine rb_vv_dvdv(tc, bc, k, vv, dd) bind(C, name='rb_vv_dvdv')
      real(c_double), intent(in) :: tc(3, 4), bc(3, 4)
      real(c_double), value :: k
      complex(c_double_complex), intent(out) :: vv, dd(3, 3)
      real(dp), allocatable :: P0_tet(:, :), w0_tet(:), P0_tri(:, :), w0_tri(:)
      real(dp) :: T_coord(3, 4), B_coord(3, 4), Pt(3, 24), wt(24), Pb(3, 24), wb(24)
      real(dp) :: Pt2(3, 7), wt2(7), Pb2(3, 7), wb2(7), tt(3, 3), bt3(3, 3), tn(3), bn(3), rf(3), kk
      complex(dp) :: intN, acc, accm(3, 3)
      integer :: i1, ft, fb
      T_coord = tc; B_coord = bc; kk = k
      call inttetra(P0_tet, w0_tet, 5)
      call inttri(P0_tri, w0_tri, 5)
      call linmap_tet(Pt, wt, T_coord, P0_tet, w0_tet)
      call linmap_tet(Pb, wb, B_coord, P0_tet, w0_tet)
      acc = dcmplx(0, 0)
      do i1 = 1, 24
         rf = Pt(:, i1)
         intN = singularity_subtraction_int_V_G(rf, kk, Pb, wb, B_coord)
         acc = acc + intN*wt(i1)
      end do
      vv = acc
      accm = dcmplx(0, 0)
      do ft = 1, 4
         tt = tetra_face(T_coord, ft)
         tn = tri_n_vectors(tt)
         call linmap_tri(Pt2, wt2, tt, P0_tri, w0_tri)
         do fb = 1, 4
            bt3 = tetra_face(B_coord, fb)
            bn = tri_n_vectors(bt3)
            call linmap_tri(Pb2, wb2, bt3, P0_tri, w0_tri)
            do i1 = 1, 7
               rf = Pt2(:, i1)
               intN = singularity_subtraction_int_S_G(rf, bt3, bn, kk, Pb2, wb2)
               accm(:, 1) = accm(:, 1) + tn*bn(1)*intN*wt2(i1)
               accm(:, 2) = accm(:, 2) + tn*bn(2)*intN*wt2(i1)
               accm(:, 3) = accm(:, 3) + tn*bn(3)*intN*wt2(i1)
            end do
         end do
      end do
      dd = accm
   end subroutine

   ! quadrature tables: which = 1 tetra, 2 triangle, 3 cube sample points
   subroutine rb_table(which, order, p, w, n) bind(C, name='rb_table')
      integer(c_int), value :: which, order
      real(c_double), intent(out) :: p(*), w(*)
      integer(c_int), intent(out) :: n
      real(dp), allocatable :: pp(:, :), ww(:)
      integer :: i
      if (which == 1) call inttetra(pp, ww, order)
      if (which == 2) call inttri(pp, ww, order)
      if (which == 3) then
         call sample_points_cube(pp)
         allocate (ww(size(pp, 2)))
         ww = 0
      end if
      n = size(ww)
      do i = 1, n
         p(size(pp, 1)*(i - 1) + 1:size(pp, 1)*i) = pp(:, i)
         w(i) = ww(i)
      end do
   end subroutine

   subroutine rb_bessel(n, zr, zi, csj) bind(C, name='rb_bessel')
      integer(c_int), value :: n
      real(c_double), value :: zr, zi
      complex(c_double_complex), intent(out) :: csj(0:n)
      complex(8) :: sj(0:n), sy(0:n), zz
      zz = dcmplx(zr, zi)
      call cspherebessel(n, zz, sj, sy)
      csj = sj
   end subroutine

   subroutine rb_legendre(n, x, p) bind(C, name='rb_legendre')
      integer(c_int), value :: n
      real(c_double), value :: x
      real(c_double), intent(out) :: p(n + 1)
      real(dp) :: pl(n + 1), xx
      xx = x
      call legendre2(n, xx, pl)
      p = pl
   end subroutine

   ! build_grid + build_box + tetras_in_cubes on a mesh (etopol 1-based)
   subroutine rb_geometry(nnode, ntet, coord, etopol, ints, dbls, tet_cube) bind(C, name='rb_geometry')
      integer(c_int), value :: nnode, ntet
      real(c_double), intent(in) :: coord(3, nnode)
      integer(c_int), intent(in) :: etopol(4, ntet)
      integer(c_int), intent(out) :: ints(8), tet_cube(ntet)
      real(c_double), intent(out) :: dbls(8)
      integer :: t, j
      if (allocated(mesh%coord)) deallocate (mesh%coord)
      if (allocated(mesh%etopol)) deallocate (mesh%etopol)
      if (allocated(mesh%nodes)) deallocate (mesh%nodes)
      if (allocated(mesh%etopol_box)) deallocate (mesh%etopol_box)
      if (allocated(mesh%tetras)) deallocate (mesh%tetras)
      allocate (mesh%coord(3, nnode), mesh%etopol(4, ntet))
      mesh%coord = coord
      mesh%etopol = etopol
      mesh%N_tet = ntet
      mesh%N_node = nnode
      call build_grid(mesh)
      call build_box(mesh)
      call tetras_in_cubes(mesh)
      ints = 0
      ints(1:8) = [mesh%Nx, mesh%Ny, mesh%Nz, mesh%N_cubes, mesh%N_tet_cube, mesh%Nx_cube, &
                   mesh%Ny_cube, mesh%Nz_cube]
      dbls = 0
      dbls(1:5) = [mesh%delta, mesh%box_delta, mesh%min_coord(1), mesh%min_coord(2), mesh%min_coord(3)]
      tet_cube = 0
      do t = 1, mesh%N_cubes
         do j = 1, mesh%N_tet_cube
            if (mesh%tetras(j, t) > 0) tet_cube(mesh%tetras(j, t)) = t
         end do
      end do
   end subroutine

end module ref_bind
