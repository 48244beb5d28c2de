! C-callable driver around the reference's own Fortran routines (TEST INFRASTRUCTURE).
!
! oracle/Makefile compiles the PETSc-free modules of Nischay-Pro/mfcfd
! src_mpi_primal unmodified from where they lie (parameter, data_structure,
! q_variables, split_fluxes, quadrant_fluxes, limiters, interior_fluxes, wall_fluxes,
! outer_fluxes, flux_residual, point_normals) and links them with this driver into
! oracle/_ref/libmfcfd_ref.so.  The driver only fills the reference's module state
! (data_structure.F90:13-49) from caller arrays, calls the reference routines and
! copies fields back out; it does no arithmetic of its own.
!
! Routines that need PETSc (state_update, func_delta, generate_connectivity,
! compute_force_coeffs) cannot be built here; the half-stencil lists that
! generate_connectivity would produce are passed in by the caller instead.
module ref_driver
    use iso_c_binding
    use data_structure_mod
    use q_variables_mod
    use flux_residual_mod
    use point_normals_mod
    use split_fluxes_mod
    use quadrant_fluxes_mod
    use limiters_mod
    implicit none

contains

    subroutine ref_setup(n, w, cx, cy, cleft, cright, cflag1, cflag2, cmin_dist, cnbhs, &
                         cconn, c_power, c_vl, c_inflag, c_format) bind(C, name="ref_setup")
        integer(c_int), value :: n, w, c_inflag, c_format
        real(c_double), intent(in) :: cx(n), cy(n), cmin_dist(n)
        integer(c_int), intent(in) :: cleft(n), cright(n), cflag1(n), cflag2(n), cnbhs(n)
        integer(c_int), intent(in) :: cconn(w, n)
        real(c_double), value :: c_power, c_vl
        integer :: i, k, nw, no, ni, ns

        call ref_release()
        max_points = n
        local_points = n
        ghost_points = 0
        power = c_power
        vl_const = c_vl
        interior_points_normal_flag = c_inflag
        format = c_format

        allocate(point%x(n), point%y(n), point%nx(n), point%ny(n), point%min_dist(n))
        allocate(point%left(n), point%right(n), point%flag_1(n), point%flag_2(n))
        allocate(point%nbhs(n), point%conn(n, w))
        allocate(point%prim(4, n), point%prim_old(4, n), point%flux_res(4, n), point%q(4, n))
        allocate(point%dq(2, 4, n), point%qm(2, 4, n), point%temp(3, 4, n), point%delta(n))
        allocate(point%xpos_nbhs(n), point%xneg_nbhs(n), point%ypos_nbhs(n), point%yneg_nbhs(n))
        allocate(point%xpos_conn(n, w), point%xneg_conn(n, w))
        allocate(point%ypos_conn(n, w), point%yneg_conn(n, w))

        point%conn = 0
        nw = 0; no = 0; ni = 0; ns = 0
        do i = 1, n
            point%x(i) = cx(i)
            point%y(i) = cy(i)
            point%min_dist(i) = cmin_dist(i)
            point%left(i) = cleft(i) + 1
            point%right(i) = cright(i) + 1
            point%flag_1(i) = cflag1(i)
            point%flag_2(i) = cflag2(i)
            point%nbhs(i) = cnbhs(i)
            do k = 1, cnbhs(i)
                point%conn(i, k) = cconn(k, i) + 1
            end do
            if (cflag1(i) == 0) nw = nw + 1
            if (cflag1(i) == 1) ni = ni + 1
            if (cflag1(i) == 2) no = no + 1
            if (cflag2(i) > 0) ns = ns + 1
        end do
        wall_points = nw
        outer_points = no
        interior_points = ni
        shape_points = ns
        allocate(wall_points_index(nw), outer_points_index(no))
        allocate(interior_points_index(ni), shape_points_index(ns))
        nw = 0; no = 0; ni = 0; ns = 0
        do i = 1, n
            if (cflag1(i) == 0) then
                nw = nw + 1; wall_points_index(nw) = i
            else if (cflag1(i) == 1) then
                ni = ni + 1; interior_points_index(ni) = i
            else if (cflag1(i) == 2) then
                no = no + 1; outer_points_index(no) = i
            end if
            if (cflag2(i) > 0) then
                ns = ns + 1; shape_points_index(ns) = i
            end if
        end do
    end subroutine

    subroutine ref_release() bind(C, name="ref_release")
        if (allocated(point%x)) then
            deallocate(point%x, point%y, point%nx, point%ny, point%min_dist)
            deallocate(point%left, point%right, point%flag_1, point%flag_2)
            deallocate(point%nbhs, point%conn)
            deallocate(point%prim, point%prim_old, point%flux_res, point%q)
            deallocate(point%dq, point%qm, point%temp, point%delta)
            deallocate(point%xpos_nbhs, point%xneg_nbhs, point%ypos_nbhs, point%yneg_nbhs)
            deallocate(point%xpos_conn, point%xneg_conn, point%ypos_conn, point%yneg_conn)
            deallocate(wall_points_index, outer_points_index)
            deallocate(interior_points_index, shape_points_index)
        end if
    end subroutine

    ! Half-stencil lists: counts (n) and 0-based lists, row-major [n][w]
    subroutine ref_set_split(w, xpn, xnn, ypn, ynn, xpc, xnc, ypc, ync) bind(C, name="ref_set_split")
        integer(c_int), value :: w
        integer(c_int), intent(in) :: xpn(max_points), xnn(max_points), ypn(max_points), ynn(max_points)
        integer(c_int), intent(in) :: xpc(w, max_points), xnc(w, max_points)
        integer(c_int), intent(in) :: ypc(w, max_points), ync(w, max_points)
        integer :: i, k
        point%xpos_conn = 0; point%xneg_conn = 0; point%ypos_conn = 0; point%yneg_conn = 0
        do i = 1, max_points
            point%xpos_nbhs(i) = xpn(i)
            point%xneg_nbhs(i) = xnn(i)
            point%ypos_nbhs(i) = ypn(i)
            point%yneg_nbhs(i) = ynn(i)
            do k = 1, xpn(i)
                point%xpos_conn(i, k) = xpc(k, i) + 1
            end do
            do k = 1, xnn(i)
                point%xneg_conn(i, k) = xnc(k, i) + 1
            end do
            do k = 1, ypn(i)
                point%ypos_conn(i, k) = ypc(k, i) + 1
            end do
            do k = 1, ynn(i)
                point%yneg_conn(i, k) = ync(k, i) + 1
            end do
        end do
    end subroutine

    ! Copy a field in or out.  which: 1 prim, 2 q, 3 dq, 4 qm, 5 delta, 6 nx, 7 ny,
    ! 8 flux_res, 9 temp(1:2,:,:).  Layout as the reference's (column-major) arrays.
    subroutine ref_field(which, dir, buf) bind(C, name="ref_field")
        integer(c_int), value :: which, dir
        real(c_double), intent(inout) :: buf(*)
        integer :: n, i, r, d, o
        n = max_points
        select case (which)
        case (1)
            if (dir == 0) then
                point%prim = reshape(buf(1:4*n), [4, n])
            else
                buf(1:4*n) = reshape(point%prim, [4*n])
            end if
        case (2)
            if (dir == 0) then
                point%q = reshape(buf(1:4*n), [4, n])
            else
                buf(1:4*n) = reshape(point%q, [4*n])
            end if
        case (3)
            if (dir == 0) then
                point%dq = reshape(buf(1:8*n), [2, 4, n])
            else
                buf(1:8*n) = reshape(point%dq, [8*n])
            end if
        case (4)
            if (dir == 0) then
                point%qm = reshape(buf(1:8*n), [2, 4, n])
            else
                buf(1:8*n) = reshape(point%qm, [8*n])
            end if
        case (5)
            if (dir == 0) then
                point%delta = buf(1:n)
            else
                buf(1:n) = point%delta
            end if
        case (6)
            if (dir == 0) then
                point%nx = buf(1:n)
            else
                buf(1:n) = point%nx
            end if
        case (7)
            if (dir == 0) then
                point%ny = buf(1:n)
            else
                buf(1:n) = point%ny
            end if
        case (8)
            if (dir == 0) then
                point%flux_res = reshape(buf(1:4*n), [4, n])
            else
                buf(1:4*n) = reshape(point%flux_res, [4*n])
            end if
        case (9)
            if (dir /= 0) then
                o = 0
                do i = 1, n
                    do r = 1, 4
                        do d = 1, 2
                            o = o + 1
                            buf(o) = point%temp(d, r, i)
                        end do
                    end do
                end do
            end if
        end select
    end subroutine

    subroutine ref_eval_q_variables() bind(C, name="ref_eval_q_variables")
        call eval_q_variables()
    end subroutine

    subroutine ref_eval_q_derivatives() bind(C, name="ref_eval_q_derivatives")
        call eval_q_derivatives()
    end subroutine

    subroutine ref_eval_q_inner_loop() bind(C, name="ref_eval_q_inner_loop")
        call eval_q_inner_loop()
    end subroutine

    subroutine ref_eval_update_innerloop() bind(C, name="ref_eval_update_innerloop")
        call eval_update_innerloop()
    end subroutine

    subroutine ref_cal_flux_residual() bind(C, name="ref_cal_flux_residual")
        call cal_flux_residual()
    end subroutine

    subroutine ref_compute_normals() bind(C, name="ref_compute_normals")
        call compute_normals()
    end subroutine

    subroutine ref_qtilde_to_primitive(qt, out) bind(C, name="ref_qtilde_to_primitive")
        real(c_double), intent(in) :: qt(4)
        real(c_double), intent(out) :: out(4)
        real(8) :: qq(4), u1, u2, rho, pr
        qq = qt
        call qtilde_to_primitive(qq, u1, u2, rho, pr)
        out(1) = rho; out(2) = u1; out(3) = u2; out(4) = pr
    end subroutine

    ! One kinetic split flux: kind 0 Gxp, 1 Gxn, 2 Gyp, 3 Gyn, 4 GxI, 5 GxII,
    ! 6 GxIII, 7 GxIV (split_fluxes.F90, quadrant_fluxes.F90)
    subroutine ref_flux(kind, cnx, cny, cu1, cu2, crho, cpr, g) bind(C, name="ref_flux")
        integer(c_int), value :: kind
        real(c_double), value :: cnx, cny, cu1, cu2, crho, cpr
        real(c_double), intent(out) :: g(4)
        real(8) :: gg(4), nx_, ny_, u1, u2, rho, pr
        nx_ = cnx; ny_ = cny; u1 = cu1; u2 = cu2; rho = crho; pr = cpr
        select case (kind)
        case (0); call flux_Gxp(gg, nx_, ny_, u1, u2, rho, pr)
        case (1); call flux_Gxn(gg, nx_, ny_, u1, u2, rho, pr)
        case (2); call flux_Gyp(gg, nx_, ny_, u1, u2, rho, pr)
        case (3); call flux_Gyn(gg, nx_, ny_, u1, u2, rho, pr)
        case (4); call flux_quad_GxI(gg, nx_, ny_, u1, u2, rho, pr)
        case (5); call flux_quad_GxII(gg, nx_, ny_, u1, u2, rho, pr)
        case (6); call flux_quad_GxIII(gg, nx_, ny_, u1, u2, rho, pr)
        case (7); call flux_quad_GxIV(gg, nx_, ny_, u1, u2, rho, pr)
        end select
        g = gg
    end subroutine

    ! venkat_limiter of point k (1-based) on the module's current q, qm, min_dist
    subroutine ref_venkat(k, qt, phi) bind(C, name="ref_venkat")
        integer(c_int), value :: k
        real(c_double), intent(in) :: qt(4)
        real(c_double), intent(out) :: phi(4)
        real(8) :: qq(4), pp(4)
        integer :: kk
        qq = qt
        kk = k
        call venkat_limiter(qq, pp, kk)
        phi = pp
    end subroutine

end module ref_driver
