! ISO_C_BINDING interface of libscadyn_amd.so for the scadyn Fortran host.
!
! Mirrors include/scadyn_amd.h.  The reference's calc_T
! (src/T_matrix.f90:56-72) replaces update_projections + build_G +
! compute_near_zone_interactions_const + compute_T_matrix by one call of
! scadyn_solve_bar; the host keeps params.in, the HDF5 mesh and T-matrix I/O.
module scadyn_amd_iface
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: SCADYN_OK = 0, SCADYN_ERR_ARG = 1, SCADYN_ERR_ORIENTATION = 2, &
                                SCADYN_ERR_GRID = 3, SCADYN_ERR_SINGULAR = 4, SCADYN_ERR_NOCONV = 5, &
                                SCADYN_ERR_MEMORY = 6, SCADYN_ERR_DEVICE = 7, SCADYN_ERR_UNSUPPORTED = 8

   type, bind(C) :: scadyn_mesh
      integer(c_int32_t) :: n_node, n_tet
      type(c_ptr) :: coord, etopol, param
   end type

   type, bind(C) :: scadyn_opts
      real(c_double) :: tol
      integer(c_int32_t) :: maxit, restart, M_ex, near_zone
      real(c_double) :: grid_size
      integer(c_int32_t) :: order, precision, batch, strict, device, fft_pad
      real(c_double) :: refine_tol
      integer(c_int32_t) :: ortho, fft_kernel, mgs_block, gather_kernel, basis
      integer(c_int32_t) :: reserved(3)
   end type

   interface
      subroutine scadyn_opts_default(opts) bind(C, name='scadyn_opts_default')
         import :: scadyn_opts
         type(scadyn_opts), intent(out) :: opts
      end subroutine

      function scadyn_last_error() bind(C, name='scadyn_last_error') result(p)
         import :: c_ptr
         type(c_ptr) :: p
      end function

      function scadyn_tmat_solve(mesh, opts, k, nmax, Taa, Tab, Tba, Tbb, iters) &
         bind(C, name='scadyn_tmat_solve') result(rc)
         import :: scadyn_mesh, scadyn_opts, c_double, c_int32_t, c_int, c_double_complex
         type(scadyn_mesh), intent(in) :: mesh
         type(scadyn_opts), intent(in) :: opts
         real(c_double), value :: k
         integer(c_int32_t), value :: nmax
         complex(c_double_complex), intent(out) :: Taa(*), Tab(*), Tba(*), Tbb(*)
         integer(c_int32_t), intent(out) :: iters(*)
         integer(c_int) :: rc
      end function

      ! tr_T (src/mie.f90:728-759): crs = C_ext, C_sca, C_abs, and the trace sum
      function scadyn_cross_sections(nmax, Taa, Tab, Tba, Tbb, k, crs) &
         bind(C, name='scadyn_cross_sections') result(rc)
         import :: c_double, c_int32_t, c_int, c_double_complex
         integer(c_int32_t), value :: nmax
         complex(c_double_complex), intent(in) :: Taa(*), Tab(*), Tba(*), Tbb(*)
         real(c_double), value :: k
         real(c_double), intent(out) :: crs(4)
         integer(c_int) :: rc
      end function

      ! z components of forcetorque (src/forces.f90:81-257) for an x-polarised
      ! plane wave along +z of amplitude E; a = volume-equivalent radius:
      ! out = F_z (N), T_z (N m), Q_f, Q_t
      function scadyn_force_torque_z(nmax, Taa, Tab, Tba, Tbb, k, E, a, out) &
         bind(C, name='scadyn_force_torque_z') result(rc)
         import :: c_double, c_int32_t, c_int, c_double_complex
         integer(c_int32_t), value :: nmax
         complex(c_double_complex), intent(in) :: Taa(*), Tab(*), Tba(*), Tbb(*)
         real(c_double), value :: k, E, a
         real(c_double), intent(out) :: out(4)
         integer(c_int) :: rc
      end function
   end interface

contains

   ! Message of the last failure as a Fortran string.
   function scadyn_error_message() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: s(:)
      integer :: n
      p = scadyn_last_error()
      call c_f_pointer(p, s, [4096])
      n = 0
      do while (n < 4096)
         if (s(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate (character(len=n) :: msg)
      msg = transfer(s(1:n), msg)
   end function

   ! One wavelength bar: the arguments are the reference's own arrays
   ! (mesh%coord(3,Nnode), mesh%etopol(4,N), mesh%param(N)) and mesh settings;
   ! T blocks come back scaled by i k^3, nm x nm with nm = nmax(nmax+2).
   function scadyn_solve_bar(coord, etopol, param, k, nmax, tol, maxit, restart, M_ex, near_zone, &
                             grid_size, Taa, Tab, Tba, Tbb, iters) result(rc)
      real(c_double), intent(in), target :: coord(:, :)
      integer(c_int32_t), intent(in), target :: etopol(:, :)
      complex(c_double_complex), intent(in), target :: param(:)
      real(c_double), intent(in) :: k, tol, grid_size
      integer, intent(in) :: nmax, maxit, restart, M_ex, near_zone
      complex(c_double_complex), intent(out) :: Taa(:, :), Tab(:, :), Tba(:, :), Tbb(:, :)
      integer(c_int32_t), intent(out) :: iters(:)
      integer(c_int) :: rc
      type(scadyn_mesh) :: m
      type(scadyn_opts) :: o
      m%n_node = size(coord, 2)
      m%n_tet = size(etopol, 2)
      m%coord = c_loc(coord)
      m%etopol = c_loc(etopol)
      m%param = c_loc(param)
      call scadyn_opts_default(o)
      o%tol = tol
      o%maxit = maxit
      o%restart = restart
      o%M_ex = M_ex
      o%near_zone = near_zone
      o%grid_size = grid_size
      rc = scadyn_tmat_solve(m, o, k, int(nmax, c_int32_t), Taa, Tab, Tba, Tbb, iters)
   end function

end module scadyn_amd_iface
