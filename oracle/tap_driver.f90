! C-callable driver around the reference's Tapenade-generated reverse routines
! (TEST INFRASTRUCTURE).
!
! oracle/Makefile compiles src_mpi_adjoint/parameter.F90 and the PETSc-free reverse
! modules of src_mpi_adjoint/sensitivity_files (data_structure_b, split_fluxes_b,
! quadrant_fluxes_b, limiters_b, q_variables_b, interior/wall/outer_fluxes_b,
! flux_residual_b, point_normals_b) unmodified from where they lie, plus Tapenade's
! ADFirstAidKit/adStack.c and adBuffer.f (push/pop stack), and links them with this driver into
! oracle/_tap/libmfcfd_tap.so.  The driver only fills the module state (point and its
! adjoint pointb, data_structure_b.F90:15-58) from caller arrays, calls the reverse
! routines and copies fields back out; it does no arithmetic of its own.
!
! Routines whose sources include PETSc headers (state_update_b, fpi_solver_b,
! compute_force_coeffs_b, the checkpointed driver) cannot be built here.
module tap_driver
    use iso_c_binding
    use data_structure_mod_diff
    use q_variables_mod_diff
    use flux_residual_mod_diff
    use point_normals_mod_diff
    use split_fluxes_mod_diff
    use quadrant_fluxes_mod_diff
    use limiters_mod_diff
    implicit none

contains

    subroutine tap_setup(n, w, cx, cy, cleft, cright, cflag1, cflag2, cmin_dist, cnbhs, &
                         cconn, c_power, c_vl, c_inflag, c_format) bind(C, name="tap_setup")
        integer(c_int), value :: n, w, c_inflag, c_format
        real(c_double), intent(in) :: cx(n), cy(n), cmin_dist(n)
        integer(c_int), intent(in) :: cleft(n), cright(n), cflag1(n), cflag2(n), cnbhs(n)
        integer(c_int), intent(in) :: cconn(w, n)
        real(c_double), value :: c_power, c_vl
        integer :: i, k, nw, no, ni, ns

        call tap_release()
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
        allocate(pointb%x(n), pointb%y(n), pointb%nx(n), pointb%ny(n), pointb%delta(n))
        allocate(pointb%prim(4, n), pointb%prim_old(4, n), pointb%flux_res(4, n), pointb%q(4, n))
        allocate(pointb%dq(2, 4, n), pointb%qm(2, 4, n), pointb%temp(3, 4, n))

        point%conn = 0
        point%prim = 0; point%prim_old = 0; point%flux_res = 0; point%q = 0
        point%dq = 0; point%qm = 0; point%temp = 0; point%delta = 0
        nw = 0; no = 0; ni = 0; ns = 0
        do i = 1, n
            point%x(i) = cx(i)
            point%y(i) = cy(i)
            point%nx(i) = 0
            point%ny(i) = 0
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
        call tap_zero_adjoints()
    end subroutine

    subroutine tap_release() bind(C, name="tap_release")
        if (allocated(point%x)) then
            deallocate(point%x, point%y, point%nx, point%ny, point%min_dist)
            deallocate(point%left, point%right, point%flag_1, point%flag_2)
            deallocate(point%nbhs, point%conn)
            deallocate(point%prim, point%prim_old, point%flux_res, point%q)
            deallocate(point%dq, point%qm, point%temp, point%delta)
            deallocate(point%xpos_nbhs, point%xneg_nbhs, point%ypos_nbhs, point%yneg_nbhs)
            deallocate(point%xpos_conn, point%xneg_conn, point%ypos_conn, point%yneg_conn)
            deallocate(pointb%x, pointb%y, pointb%nx, pointb%ny, pointb%delta)
            deallocate(pointb%prim, pointb%prim_old, pointb%flux_res, pointb%q)
            deallocate(pointb%dq, pointb%qm, pointb%temp)
            deallocate(wall_points_index, outer_points_index)
            deallocate(interior_points_index, shape_points_index)
        end if
    end subroutine

    subroutine tap_zero_adjoints() bind(C, name="tap_zero_adjoints")
        pointb%x = 0; pointb%y = 0; pointb%nx = 0; pointb%ny = 0; pointb%delta = 0
        pointb%prim = 0; pointb%prim_old = 0; pointb%flux_res = 0; pointb%q = 0
        pointb%dq = 0; pointb%qm = 0; pointb%temp = 0
    end subroutine

    ! Half-stencil lists: counts (n) and 0-based lists, row-major [n][w]
    subroutine tap_set_split(w, xpn, xnn, ypn, ynn, xpc, xnc, ypc, ync) bind(C, name="tap_set_split")
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

    ! Copy one field of point (adj == 0) or of its adjoint pointb (adj /= 0) in
    ! (dir == 0) or out.  which: 1 prim, 2 q, 3 dq, 4 qm, 5 delta, 6 nx, 7 ny,
    ! 8 flux_res, 10 x, 11 y.  Layouts as the reference's (column-major) arrays.
    subroutine tap_field(which, adj, dir, buf) bind(C, name="tap_field")
        integer(c_int), value :: which, adj, dir
        real(c_double), intent(inout) :: buf(*)
        integer :: n
        n = max_points
        if (adj == 0) then
            select case (which)
            case (1); call io2(point%prim)
            case (2); call io2(point%q)
            case (3); call io3(point%dq)
            case (4); call io3(point%qm)
            case (5); call io1(point%delta)
            case (6); call io1(point%nx)
            case (7); call io1(point%ny)
            case (8); call io2(point%flux_res)
            case (10); call io1(point%x)
            case (11); call io1(point%y)
            end select
        else
            select case (which)
            case (1); call io2(pointb%prim)
            case (2); call io2(pointb%q)
            case (3); call io3(pointb%dq)
            case (4); call io3(pointb%qm)
            case (5); call io1(pointb%delta)
            case (6); call io1(pointb%nx)
            case (7); call io1(pointb%ny)
            case (8); call io2(pointb%flux_res)
            case (10); call io1(pointb%x)
            case (11); call io1(pointb%y)
            end select
        end if
    contains
        subroutine io1(a)
            real(8), intent(inout) :: a(:)
            if (dir == 0) then
                a = buf(1:n)
            else
                buf(1:n) = a
            end if
        end subroutine
        subroutine io2(a)
            real(8), intent(inout) :: a(:, :)
            if (dir == 0) then
                a = reshape(buf(1:4*n), [4, n])
            else
                buf(1:4*n) = reshape(a, [4*n])
            end if
        end subroutine
        subroutine io3(a)
            real(8), intent(inout) :: a(:, :, :)
            if (dir == 0) then
                a = reshape(buf(1:8*n), [2, 4, n])
            else
                buf(1:8*n) = reshape(a, [8*n])
            end if
        end subroutine
    end subroutine

    ! forward routines of the differentiated modules (same arithmetic as src_mpi_primal)
    subroutine tap_compute_normals() bind(C, name="tap_compute_normals")
        call compute_normals()
    end subroutine
    subroutine tap_eval_q_variables() bind(C, name="tap_eval_q_variables")
        call eval_q_variables()
    end subroutine
    subroutine tap_eval_q_derivatives() bind(C, name="tap_eval_q_derivatives")
        call eval_q_derivatives()
    end subroutine
    subroutine tap_cal_flux_residual() bind(C, name="tap_cal_flux_residual")
        call cal_flux_residual()
    end subroutine

    ! reverse routines: adjoints of the outputs in pointb in, adjoints of the inputs
    ! added into pointb
    subroutine tap_compute_normals_b() bind(C, name="tap_compute_normals_b")
        call compute_normals_b()
    end subroutine
    subroutine tap_eval_q_variables_b() bind(C, name="tap_eval_q_variables_b")
        call eval_q_variables_b()
    end subroutine
    subroutine tap_eval_q_derivatives_b() bind(C, name="tap_eval_q_derivatives_b")
        call eval_q_derivatives_b()
    end subroutine
    subroutine tap_eval_q_inner_loop_b() bind(C, name="tap_eval_q_inner_loop_b")
        call eval_q_inner_loop_b()
    end subroutine
    subroutine tap_eval_update_innerloop_b() bind(C, name="tap_eval_update_innerloop_b")
        call eval_update_innerloop_b()
    end subroutine
    subroutine tap_cal_flux_residual_b() bind(C, name="tap_cal_flux_residual_b")
        call cal_flux_residual_b()
    end subroutine

    ! Reverse of one kinetic flux (kind 0 Gxp, 1 Gxn, 2 Gyp, 3 Gyn, 4 GxI, 5 GxII,
    ! 6 GxIII, 7 GxIV) at (nx, ny, u1, u2, rho, pr) with output adjoint gb:
    ! out = (nxb, nyb, u1b, u2b, rhob, prb)
    subroutine tap_flux_b(kind, cnx, cny, cu1, cu2, crho, cpr, gb, out) bind(C, name="tap_flux_b")
        integer(c_int), value :: kind
        real(c_double), value :: cnx, cny, cu1, cu2, crho, cpr
        real(c_double), intent(in) :: gb(4)
        real(c_double), intent(out) :: out(6)
        real(8) :: g(4), gbb(4), nx_, ny_, u1, u2, rho, pr
        real(8) :: nxb, nyb, u1b, u2b, rhob, prb
        nx_ = cnx; ny_ = cny; u1 = cu1; u2 = cu2; rho = crho; pr = cpr
        nxb = 0; nyb = 0; u1b = 0; u2b = 0; rhob = 0; prb = 0
        g = 0
        gbb = gb
        select case (kind)
        case (0); call flux_gxp_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        case (1); call flux_gxn_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        case (2); call flux_gyp_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        case (3); call flux_gyn_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        case (4); call flux_quad_gxi_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        case (5); call flux_quad_gxii_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        case (6); call flux_quad_gxiii_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        case (7); call flux_quad_gxiv_b(g, gbb, nx_, nxb, ny_, nyb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        end select
        out(1) = nxb; out(2) = nyb; out(3) = u1b; out(4) = u2b; out(5) = rhob; out(6) = prb
    end subroutine

    ! qtilde_to_primitive of the differentiated module: out = (rho, u1, u2, pr)
    subroutine tap_qtilde_to_primitive(qt, out) bind(C, name="tap_qtilde_to_primitive")
        real(c_double), intent(in) :: qt(4)
        real(c_double), intent(out) :: out(4)
        real(8) :: qq(4), u1, u2, rho, pr
        qq = qt
        call qtilde_to_primitive(qq, u1, u2, rho, pr)
        out(1) = rho; out(2) = u1; out(3) = u2; out(4) = pr
    end subroutine

    ! Reverse of qtilde_to_primitive: pb = (rhob, u1b, u2b, prb) -> qtb
    subroutine tap_qtilde_to_primitive_b(qt, pb, qtb) bind(C, name="tap_qtilde_to_primitive_b")
        real(c_double), intent(in) :: qt(4), pb(4)
        real(c_double), intent(out) :: qtb(4)
        real(8) :: qq(4), qqb(4), u1, u2, rho, pr, u1b, u2b, rhob, prb
        qq = qt
        u1 = 0; u2 = 0; rho = 0; pr = 0
        rhob = pb(1); u1b = pb(2); u2b = pb(3); prb = pb(4)
        qqb = 0
        call qtilde_to_primitive_b(qq, qqb, u1, u1b, u2, u2b, rho, rhob, pr, prb)
        qtb = qqb
    end subroutine

    ! Reverse of venkat_limiter for point k (1-based) on the module's q, qm, min_dist:
    ! phib -> qtb; the adjoints of q(:, k) and qm(:, :, k) are added into pointb
    subroutine tap_venkat_b(k, qt, phib, qtb) bind(C, name="tap_venkat_b")
        integer(c_int), value :: k
        real(c_double), intent(in) :: qt(4), phib(4)
        real(c_double), intent(out) :: qtb(4)
        real(8) :: qq(4), qqb(4), pp(4), ppb(4)
        integer :: kk
        qq = qt
        ppb = phib
        pp = 0
        qqb = 0
        kk = k
        call venkat_limiter_b(qq, qqb, pp, ppb, kk)
        qtb = qqb
    end subroutine

end module tap_driver
