! ISO_C_BINDING interfaces of include/mfcfd.h for a Fortran host.
!
! The reference drives its hot path through argument-less module procedures over
! module state (src_mpi_primal data_structure.F90:13-49); a host that keeps the
! reference's I/O (readcase, point reader, writers) calls these functions instead.
module mfcfd_api
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: MF_OK = 0, MF_ERR_HIP = -1, MF_ERR_INVALID = -2, MF_ERR_STATE = -3
    integer(c_int32_t), parameter :: MF_FP64 = 0, MF_MIXED = 1
    ! mf_tuning: staged walks (MF_LPP_STAGED + lanes: 1, 2, 4 for LSQ and inner, 2, 4, 8 for flux);
    ! resident inner sweeps (MF_LPP_RESIDENT + 1, 2 or 4, mixed contexts)
    integer(c_int32_t), parameter :: MF_LPP_STAGED = 16, MF_LPP_RESIDENT = 32
    ! mf_adjoint*: snaps value that tapes every iteration in the primal sweep
    integer(c_int32_t), parameter :: MF_ADJ_TAPE_ALL = -1
    integer(c_int32_t), parameter :: MF_FIELD_Q = 1, MF_FIELD_DQ = 2, MF_FIELD_QM = 3, &
        MF_FIELD_FLUX_RES = 4, MF_FIELD_DELTA = 5, MF_FIELD_PRIM_OLD = 6, MF_FIELD_NX = 7, &
        MF_FIELD_NY = 8

    type, bind(C) :: mf_geometry
        integer(c_int32_t) :: n_points, max_nbhs
        type(c_ptr) :: x, y, nx, ny, min_dist, left, right, flag_1, flag_2, nbhs, conn
    end type

    type, bind(C) :: mf_params
        real(c_double) :: cfl, vl_const, power, mach, aoa, euler, res_old
        integer(c_int32_t) :: rks, inner_iterations, shapes, interior_points_normal_flag
        integer(c_int32_t) :: format, restart, precision, reorder, block_size, first_order
    end type

    type, bind(C) :: mf_tuning
        integer(c_int32_t) :: lpp_lsq, lpp_inner, lpp_flux, barrier_polls
        real(c_double) :: us_lsq, us_inner, us_flux
        integer(c_int32_t) :: lpp_flux_rev, lpp_lsq_rev, lpp_inner_rev, lpp_gather
    end type

    type, bind(C) :: mf_adjoint_stats
        integer(c_int64_t) :: forward_steps, taped_steps
        integer(c_int32_t) :: snapshots, reserved
        real(c_double) :: primal_ms, adjoint_ms
        integer(c_int32_t) :: lpp_flux_rev, lpp_lsq_rev, lpp_inner_rev, lpp_gather
    end type

    interface
        subroutine mf_default_params(p) bind(C, name="mf_default_params")
            import
            type(mf_params), intent(out) :: p
        end subroutine
        integer(c_int) function mf_create(g, p, ctx) bind(C, name="mf_create")
            import
            type(mf_geometry), intent(in) :: g
            type(mf_params), intent(in) :: p
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function mf_destroy(ctx) bind(C, name="mf_destroy")
            import
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function mf_last_error() bind(C, name="mf_last_error")
            import
        end function
        integer(c_int) function mf_set_prim(ctx, prim) bind(C, name="mf_set_prim")
            import
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: prim(4, *)
        end function
        integer(c_int) function mf_get_prim(ctx, prim) bind(C, name="mf_get_prim")
            import
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: prim(4, *)
        end function
        integer(c_int) function mf_get_field(ctx, field, out) bind(C, name="mf_get_field")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function mf_set_field(ctx, field, in) bind(C, name="mf_set_field")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field
            real(c_double), intent(in) :: in(*)
        end function
        integer(c_int) function mf_func_delta(ctx) bind(C, name="mf_func_delta")
            import
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mf_eval_q_variables(ctx) bind(C, name="mf_eval_q_variables")
            import
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mf_eval_q_derivatives(ctx) bind(C, name="mf_eval_q_derivatives")
            import
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mf_eval_q_inner_loop(ctx) bind(C, name="mf_eval_q_inner_loop")
            import
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mf_cal_flux_residual(ctx) bind(C, name="mf_cal_flux_residual")
            import
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mf_state_update(ctx, rk, sum_res_sqr) bind(C, name="mf_state_update")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: rk
            real(c_double), intent(out) :: sum_res_sqr
        end function
        integer(c_int) function mf_compute_cl_cd_cm(ctx, cl, cd, cm, cp) bind(C, name="mf_compute_cl_cd_cm")
            import
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: cl(*), cd(*), cm(*), cp(*)
        end function
        integer(c_int32_t) function mf_shape_points(ctx) bind(C, name="mf_shape_points")
            import
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mf_run(ctx, first_it, iters, residue, res_new, cd, done) bind(C, name="mf_run")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: first_it, iters
            real(c_double), intent(out) :: residue(*), res_new(*), cd(*)
            integer(c_int32_t), intent(out) :: done
        end function
        integer(c_int) function mf_run_async(ctx, first_it, iters) bind(C, name="mf_run_async")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: first_it, iters
        end function
        integer(c_int) function mf_sync(ctx) bind(C, name="mf_sync")
            import
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function mf_autotune(ctx, reps, t) bind(C, name="mf_autotune")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: reps
            type(mf_tuning), intent(out) :: t
        end function
        integer(c_int) function mf_get_tuning(ctx, t) bind(C, name="mf_get_tuning")
            import
            type(c_ptr), value :: ctx
            type(mf_tuning), intent(out) :: t
        end function
        integer(c_int) function mf_set_tuning(ctx, t) bind(C, name="mf_set_tuning")
            import
            type(c_ptr), value :: ctx
            type(mf_tuning), intent(in) :: t
        end function
        integer(c_int) function mf_adjoint(ctx, first_it, iters, snaps, J, dJdx, dJdy, stats) &
                bind(C, name="mf_adjoint")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: first_it, iters, snaps
            real(c_double), intent(out) :: J, dJdx(*), dJdy(*)
            type(mf_adjoint_stats), intent(out) :: stats
        end function
        ! forward mode (src_mpi_forward q_lskum_d): J and its derivative along (vx, vy)
        integer(c_int) function mf_tangent(ctx, iters, vx, vy, J, dJ, prim_d) bind(C, name="mf_tangent")
            import
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: iters
            real(c_double), intent(in) :: vx(*), vy(*)
            real(c_double), intent(out) :: J, dJ
            type(c_ptr), value :: prim_d
        end function
        ! force coefficients from device arrays, no allocation (no reference counterpart)
        integer(c_int) function mf_compute_cl_cd_cm_noalloc(n_points, n_shape, shapes, mach, aoa, x, y, &
                nx, ny, prim, left, right, flag_2, shape_idx, clcdcm, cp, stream) &
                bind(C, name="mf_compute_cl_cd_cm_noalloc")
            import
            integer(c_int32_t), value :: n_points, n_shape, shapes
            real(c_double), value :: mach, aoa
            type(c_ptr), value :: x, y, nx, ny, prim, left, right, flag_2, shape_idx, clcdcm, cp, stream
        end function
    end interface

contains

    ! mf_last_error() as a Fortran string
    function mf_error_message() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: cs(:)
        integer :: n
        call c_f_pointer(mf_last_error(), cs, [1024])
        n = 0
        do while (n < 1024)
            if (cs(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(cs(1:n), msg)
    end function

end module mfcfd_api
