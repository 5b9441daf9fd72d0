! Fortran host driver for one wavelength bar through scadyn_amd_iface: the call
! sequence the reference's calc_T makes after the overlay.  Reads a mesh in the
! reference's in-memory layout from a raw file, solves, writes the four T
! blocks (column-major complex(8)) and the status, and prints the bar's
! orientation-averaged cross sections.
!
! input file (native endianness): int32 n_node, n_tet, nmax;
!   real(8) k; coord(3,n_node); int32 etopol(4,n_tet); complex(8) param(n_tet)
! output file: int32 rc; int32 iters(2 nm); complex(8) Taa,Tab,Tba,Tbb(nm,nm)
program solve_bar_main
   use iso_c_binding
   use scadyn_amd_iface
   implicit none
   character(len=1024) :: fin, fout
   integer(c_int32_t) :: n_node, n_tet, nmax, rc
   real(c_double) :: k
   real(c_double), allocatable :: coord(:, :)
   integer(c_int32_t), allocatable :: etopol(:, :), iters(:)
   complex(c_double_complex), allocatable :: param(:), Taa(:, :), Tab(:, :), Tba(:, :), Tbb(:, :)
   real(c_double) :: crs(4)
   integer :: nm, u
   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read (u) n_node, n_tet, nmax
   read (u) k
   allocate (coord(3, n_node), etopol(4, n_tet), param(n_tet))
   read (u) coord
   read (u) etopol
   read (u) param
   close (u)
   nm = (nmax + 1)**2 - 1
   allocate (Taa(nm, nm), Tab(nm, nm), Tba(nm, nm), Tbb(nm, nm), iters(2*nm))
   Taa = 0; Tab = 0; Tba = 0; Tbb = 0; iters = 0
   ! the reference defaults (src/common.f90:80-89)
   rc = scadyn_solve_bar(coord, etopol, param, k, int(nmax), 1d-5, 50, 4, 2, 1, 1.8d0, &
                         Taa, Tab, Tba, Tbb, iters)
   if (rc /= SCADYN_OK) print *, 'scadyn error: ', scadyn_error_message()
   ! the first consumer of the bar's T: orientation-averaged cross sections
   if (rc == SCADYN_OK) then
      if (scadyn_cross_sections(nmax, Taa, Tab, Tba, Tbb, k, crs) == SCADYN_OK) &
         write (*, '(a, 3es24.16)') 'cross sections ', crs(1:3)
   end if
   open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (u) rc
   write (u) iters
   write (u) Taa, Tab, Tba, Tbb
   close (u)
end program solve_bar_main
