! C-callable driver around the reference libDelhommeau Fortran modules.
!
! Test infrastructure only: this file is compiled together with the
! reference sources found under libs/libDelhommeau/src of the reference tree
! (constants.f90, float64.f90, Delhommeau_integrals.f90, Green_Rankine.f90,
! FinGreen3D.f90, Green_wave.f90, matrices.f90) into
! oracle/_ref/libdelhommeau_ref.so.  The
! resulting library is used to write golden vectors and to pin the C oracle.
! It is never part of the product path.
!
! Every argument is an explicit-shape array in Fortran (column-major) order,
! so that the caller can pass numpy arrays in "F" order by pointer.
module ref_driver
  use iso_c_binding
  use floating_point_precision, only: pre
  use constants
  use delhommeau_integrals
  use green_rankine
  use green_wave
  use matrices
  use fingreen3d_module
  implicit none
contains

  subroutine ref_default_r_spacing(nr, rmax, method, out) bind(C, name="ref_default_r_spacing")
    integer(c_int), value :: nr, method
    real(c_double), value :: rmax
    real(c_double), intent(out) :: out(nr)
    out = default_r_spacing(nr, rmax, method)
  end subroutine

  subroutine ref_default_z_spacing(nz, zmin, method, out) bind(C, name="ref_default_z_spacing")
    integer(c_int), value :: nz, method
    real(c_double), value :: zmin
    real(c_double), intent(out) :: out(nz)
    out = default_z_spacing(nz, zmin, method)
  end subroutine

  subroutine ref_construct_tabulation(nr, r_range, nz, z_range, npts, tab) &
      bind(C, name="ref_construct_tabulation")
    integer(c_int), value :: nr, nz, npts
    real(c_double), intent(in) :: r_range(nr), z_range(nz)
    real(c_double), intent(out) :: tab(nr, nz, nb_tabulated_values)
    tab = construct_tabulation(r_range, z_range, npts)
  end subroutine

  subroutine ref_numerical_integration(r, z, npts, out) bind(C, name="ref_numerical_integration")
    real(c_double), value :: r, z
    integer(c_int), value :: npts
    real(c_double), intent(out) :: out(nb_tabulated_values)
    out = numerical_integration(r, z, npts)
  end subroutine

  subroutine ref_asymptotic_approximations(r, z, out) bind(C, name="ref_asymptotic_approximations")
    real(c_double), value :: r, z
    real(c_double), intent(out) :: out(nb_tabulated_values)
    out = asymptotic_approximations(r, z)
  end subroutine

  subroutine ref_pick_in_default_tabulation(np, r, z, method, nr, r_range, nz, z_range, tab, out) &
      bind(C, name="ref_pick_in_default_tabulation")
    integer(c_int), value :: np, method, nr, nz
    real(c_double), intent(in) :: r(np), z(np)
    real(c_double), intent(in) :: r_range(nr), z_range(nz), tab(nr, nz, nb_tabulated_values)
    real(c_double), intent(out) :: out(nb_tabulated_values, np)
    integer :: i
    do i = 1, np
      out(:, i) = pick_in_default_tabulation(r(i), z(i), method, r_range, z_range, tab)
    end do
  end subroutine

  subroutine ref_nearest_indices(np, r, z, method, nr, r_range, nz, z_range, ir, iz) &
      bind(C, name="ref_nearest_indices")
    integer(c_int), value :: np, method, nr, nz
    real(c_double), intent(in) :: r(np), z(np), r_range(nr), z_range(nz)
    integer(c_int), intent(out) :: ir(np), iz(np)
    integer :: i
    do i = 1, np
      ir(i) = nearest_r_index(r(i), r_range, method)
      iz(i) = nearest_z_index(z(i), z_range, method)
    end do
  end subroutine

  subroutine ref_wave_part_infinite_depth(np, x0i, x0j, k, npts, grid, nr, r_range, nz, z_range, tab, &
      gfs, g, nablag) bind(C, name="ref_wave_part_infinite_depth")
    integer(c_int), value :: np, npts, grid, nr, nz, gfs
    real(c_double), value :: k
    real(c_double), intent(in) :: x0i(np, 3), x0j(np, 3)
    real(c_double), intent(in) :: r_range(nr), z_range(nz), tab(nr, nz, nb_tabulated_values)
    complex(c_double_complex), intent(out) :: g(np), nablag(np, 3)
    integer :: i
    complex(kind=pre) :: gg, ng(3)
    do i = 1, np
      call wave_part_infinite_depth(x0i(i, :), x0j(i, :), k, npts, grid, r_range, z_range, tab, gfs, gg, ng)
      g(i) = gg
      nablag(i, :) = ng
    end do
  end subroutine

  ! kind: 0 = integral_of_Rankine, 1 = exact, 2 = one point,
  !       3 = integral_of_reflected_Rankine, 4 = exact reflected, 5 = one point reflected
  subroutine ref_rankine(np, kind_, m, nodes, centers, normals, areas, radii, deriv_first, coefs, &
      int_g, int_nabla_g) bind(C, name="ref_rankine")
    integer(c_int), value :: np, kind_, deriv_first
    real(c_double), intent(in) :: m(np, 3), nodes(np, 4, 3), centers(np, 3), normals(np, 3)
    real(c_double), intent(in) :: areas(np), radii(np), coefs(2)
    real(c_double), intent(out) :: int_g(np), int_nabla_g(np, 3)
    integer :: i
    logical :: d
    real(kind=pre) :: g, ng(3), fn(4, 3)
    d = (deriv_first /= 0)
    do i = 1, np
      fn = nodes(i, :, :)
      select case (kind_)
      case (0)
        call integral_of_Rankine(m(i, :), fn, centers(i, :), normals(i, :), areas(i), radii(i), d, g, ng)
      case (1)
        call exact_integral_of_Rankine(m(i, :), fn, centers(i, :), normals(i, :), areas(i), radii(i), d, g, ng)
      case (2)
        call one_point_integral_of_Rankine(m(i, :), centers(i, :), areas(i), d, g, ng)
      case (3)
        call integral_of_reflected_Rankine(m(i, :), fn, centers(i, :), normals(i, :), areas(i), radii(i), &
          d, coefs, g, ng)
      case (4)
        call exact_integral_of_reflected_Rankine(m(i, :), fn, centers(i, :), normals(i, :), areas(i), &
          radii(i), d, coefs, g, ng)
      case default
        call one_point_integral_of_reflected_Rankine(m(i, :), centers(i, :), areas(i), d, coefs, g, ng)
      end select
      int_g(i) = g
      int_nabla_g(i, :) = ng
    end do
  end subroutine

  subroutine ref_build_matrices(nc, colloc, ndp, dpn, nv, nf, vertices, faces1, centers, normals, &
      areas, radii, nq, qp, qw, k, depth, npts, grid, nr, r_range, nz, z_range, tab, &
      fdm, nexp, prony, gfs, adjoint, nk, s, kk) bind(C, name="ref_build_matrices")
    integer(c_int), value :: nc, ndp, nv, nf, nq, npts, grid, nr, nz, fdm, nexp, gfs, adjoint, nk
    real(c_double), value :: k, depth
    real(c_double), intent(in) :: colloc(nc, 3), dpn(ndp, 3), vertices(nv, 3)
    integer(c_int), intent(in) :: faces1(nf, 4)
    real(c_double), intent(in) :: centers(nf, 3), normals(nf, 3), areas(nf), radii(nf)
    real(c_double), intent(in) :: qp(nf, nq, 3), qw(nf, nq)
    real(c_double), intent(in) :: r_range(nr), z_range(nz), tab(nr, nz, nb_tabulated_values)
    real(c_double), intent(in) :: prony(2, nexp)
    complex(c_double_complex), intent(inout) :: s(nc, nf), kk(nk, nc, nf)
    real(kind=pre) :: roots(1)
    integer :: ifaces(nf, 4)
    roots = 0
    ifaces = faces1
    call build_matrices(nc, colloc, dpn, nv, nf, vertices, ifaces, centers, normals, areas, radii, &
      nq, qp, qw, k, depth, npts, grid, r_range, z_range, tab, fdm, prony, roots, gfs, (adjoint /= 0), s, kk)
  end subroutine

  subroutine ref_add_diagonal_term(n, centers, ndp, dpn, free_surface, nk, nc, nf, kk) &
      bind(C, name="ref_add_diagonal_term")
    integer(c_int), value :: n, ndp, nk, nc, nf
    real(c_double), value :: free_surface
    real(c_double), intent(in) :: centers(n, 3), dpn(ndp, 3)
    complex(c_double_complex), intent(inout) :: kk(nk, nc, nf)
    call add_diagonal_term(centers, dpn, free_surface, kk)
  end subroutine

  ! ---- FinGreen3D (HAMS finite-depth Green function) ---------------------
  ! The pointwise routine (FinGreen3D.f90:71): rr, zf, zp per point, v the
  ! deep-water wavenumber omega^2/g, wvn the nk roots of the dispersion
  ! relation (wvn(1) the propagating one), h the depth.
  subroutine ref_fingreen3d(np, rr, zf, zp, v, nk, wvn, h, grn) bind(C, name="ref_fingreen3d")
    integer(c_int), value :: np, nk
    real(c_double), value :: v, h
    real(c_double), intent(in) :: rr(np), zf(np), zp(np), wvn(nk)
    complex(c_double_complex), intent(out) :: grn(np, 3)
    integer :: i
    complex(kind=8) :: g3(3)
    do i = 1, np
      call fingreen3d_routine(rr(i), zf(i), zp(i), v, wvn, nk, h, g3)
      grn(i, :) = g3
    end do
  end subroutine

  ! Newman's iteration for the roots (FinGreen3D.f90:156), as called by
  ! compute_dispersion_roots (Green_wave.f90:523-541).
  subroutine ref_fg_dispersion_roots(nk, omega2_over_g, depth, out) bind(C, name="ref_fg_dispersion_roots")
    integer(c_int), value :: nk
    real(c_double), value :: omega2_over_g, depth
    real(c_double), intent(out) :: out(nk)
    out = compute_dispersion_roots(nk, omega2_over_g, depth)
  end subroutine

  ! The special functions the series use, pointwise.
  ! kind: 0 = IK01A (K0, K1), 1 = JY01B (J0, J1, Y0, Y1), 2 = E1XA,
  !       3 = EIX, 4 = ERROR, 5 = ENXA (order n)
  subroutine ref_fg_special(np, kind_, n, x, out) bind(C, name="ref_fg_special")
    integer(c_int), value :: np, kind_, n
    real(c_double), intent(in) :: x(np)
    real(c_double), intent(out) :: out(4, np)
    integer :: i
    out = 0
    do i = 1, np
      select case (kind_)
      case (0)
        call IK01A(x(i), out(1, i), out(2, i))
      case (1)
        call JY01B(x(i), out(1, i), out(2, i), out(3, i), out(4, i))
      case (2)
        call E1XA(x(i), out(1, i))
      case (3)
        call EIX(x(i), out(1, i))
      case (4)
        call ERROR(x(i), out(1, i))
      case default
        call ENXA(n, x(i), out(1, i))
      end select
    end do
  end subroutine

  ! The epsilon-algorithm limit of the partial sums s(0:n) (FinGreen3D.f90:3053).
  subroutine ref_fg_limes(n, s, out) bind(C, name="ref_fg_limes")
    integer(c_int), value :: n
    real(c_double), intent(in) :: s(0:n)
    real(c_double), intent(out) :: out(2)
    real(c_double) :: work(0:n), lim, err
    integer :: ip, kp
    work = s
    call Limes(n, work, lim, ip, kp, err)
    out(1) = lim
    out(2) = err
  end subroutine

  ! build_matrices with the dispersion roots argument (FinGreen3D method).
  subroutine ref_build_matrices_roots(nc, colloc, ndp, dpn, nv, nf, vertices, faces1, centers, normals, &
      areas, radii, nq, qp, qw, k, depth, fdm, nroots, roots, gfs, adjoint, nk, s, kk) &
      bind(C, name="ref_build_matrices_roots")
    integer(c_int), value :: nc, ndp, nv, nf, nq, fdm, nroots, gfs, adjoint, nk
    real(c_double), value :: k, depth
    real(c_double), intent(in) :: colloc(nc, 3), dpn(ndp, 3), vertices(nv, 3)
    integer(c_int), intent(in) :: faces1(nf, 4)
    real(c_double), intent(in) :: centers(nf, 3), normals(nf, 3), areas(nf), radii(nf)
    real(c_double), intent(in) :: qp(nf, nq, 3), qw(nf, nq), roots(nroots)
    complex(c_double_complex), intent(inout) :: s(nc, nf), kk(nk, nc, nf)
    real(kind=pre) :: r_range(2), z_range(2), tab(2, 2, nb_tabulated_values), prony(2, 1)
    integer :: ifaces(nf, 4)
    r_range = 0
    z_range = 0
    tab = 0
    prony = 0
    ifaces = faces1
    call build_matrices(nc, colloc, dpn, nv, nf, vertices, ifaces, centers, normals, areas, radii, &
      nq, qp, qw, k, depth, 0, 0, r_range, z_range, tab, fdm, prony, roots, gfs, (adjoint /= 0), s, kk)
  end subroutine

end module ref_driver
