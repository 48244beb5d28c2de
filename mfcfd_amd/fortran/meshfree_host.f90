! The meshfree solver's host program over the MI355X C ABI (include/mfcfd.h).
!
! Same run as the reference's main program (src_mpi_primal/meshfree_solver.F90:1-104
! with q_lskum.F90:15-78), one GPU instead of MPI ranks: read case.in, read the point
! file, set the freestream or restart state, iterate, and write the residue history,
! cp/cp-file, solution/sol.dat and (with -adjoint_mode checkpoints -chkpts S, as
! src_mpi_adjoint/meshfree_solver.F90 runs q_lskum_chkpts_b, or -adjoint_mode tape to
! tape every iteration in the primal sweep) the Cd sensitivities in
! sensitivity/sensitivity.dat; with -adjoint_mode forward, src_mpi_forward's tangent of
! Cd along x of point 1.  Normals, connectivity splits, func_delta, the stages
! and the force coefficients all run on the device.
!
! Usage: run from a case directory holding case.in and point/point.h5 (or
! point/partGrid).  "meshfree_host --check" reads and validates the inputs, prints
! their summary and exits without touching the GPU.
!
! Without a case.in, an input.nml selects the CUDA build's run (src_cuda
! meshfree_solver.F90:60-110): its namelist options, restart from restart.dat, and
! output.dat (primitive state, Mach, entropy, adaptation sensor) at every nsave
! iterations and at the end instead of solution/sol.dat.
program meshfree_host
    use iso_c_binding
    use mfcfd_api
    use host_io
    implicit none

    type(case_options) :: opt
    type(point_data), target :: pts
    type(mf_geometry) :: geo
    type(mf_params) :: prm
    type(mf_tuning) :: tun
    type(mf_adjoint_stats) :: ast
    type(c_ptr) :: ctx
    real(c_double), allocatable, target :: prim(:, :), res(:), rnew(:), cdh(:), dJdx(:), dJdy(:)
    real(c_double), allocatable :: cl(:), cd(:), cm(:), cp(:)
    real(c_double), allocatable, target :: fx(:), fy(:)
    integer(c_int32_t), allocatable, target :: left(:), right(:), flag_1(:), flag_2(:), nbhs(:), conn(:, :)
    real(c_double) :: J, Jd
    integer :: itr, it, last, chunk, done, k, nshape, maxchunk, nw, no, ni
    integer(c_int) :: rc
    integer(8) :: c0, c1, c2, rate
    real(8) :: res_old, runtime, totaltime
    character(len=32) :: arg
    logical :: check_only

    check_only = .false.
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        check_only = (trim(arg) == '--check')
    end if

    call system_clock(c0, rate)
    write(*, *)
    write(*, *) '%%%%%%%%%%%%%-MI355X Meshfree Code-%%%%%%%%%%%'
    write(*, *)
    write(*, *) '%%%%%%%%%%%-Reading the case file-%%%%%%%%%'
    write(*, *)
    if (file_exists('case.in')) then
        call read_case('case.in', opt)
    else if (file_exists('input.nml')) then
        call read_nml('input.nml', opt)
    else
        write(*, '(a)') ' *** neither case.in nor input.nml in the working directory'
        error stop 2
    end if
    call print_options(opt)
    write(*, *)
    write(*, *) '%%%%%%%%%%%%-Reading point data-%%%%%%%%%%%'
    write(*, *)
    call read_points(opt, pts)
    nw = count(pts%flag_1 == 0)
    ni = count(pts%flag_1 == 1)
    no = count(pts%flag_1 == 2)
    nshape = count(pts%flag_2 > 0)
    write(*, *) 'Number of points:         ', pts%n
    write(*, *) 'wall, interior, outer, shape:', nw, ni, no, nshape
    write(*, *) 'max neighbours:           ', pts%w
    if (check_only) then
        write(*, '(a,i0,a,i0,a,i0,a,es24.16)') 'check: n=', pts%n, ' w=', pts%w, ' sum_nbhs=', sum(pts%nbhs), &
            ' sum_xy=', sum(pts%x) + sum(pts%y)
        stop
    end if
    call execute_command_line('mkdir -p solution cp')
    if (opt%ad_mode == 1) call execute_command_line('mkdir -p sensitivity')

    ! device view of the point data (conn(N, W) column-major, 1-based)
    allocate(fx(pts%n), fy(pts%n), left(pts%n), right(pts%n), flag_1(pts%n), flag_2(pts%n), nbhs(pts%n))
    allocate(conn(pts%n, pts%w))
    fx = pts%nx
    fy = pts%ny
    left = pts%left
    right = pts%right
    flag_1 = pts%flag_1
    flag_2 = pts%flag_2
    nbhs = pts%nbhs
    conn = pts%conn
    geo%n_points = pts%n
    geo%max_nbhs = pts%w
    geo%x = c_loc(pts%x)
    geo%y = c_loc(pts%y)
    geo%nx = c_loc(fx)
    geo%ny = c_loc(fy)
    geo%min_dist = c_loc(pts%min_dist)
    geo%left = c_loc(left)
    geo%right = c_loc(right)
    geo%flag_1 = c_loc(flag_1)
    geo%flag_2 = c_loc(flag_2)
    geo%nbhs = c_loc(nbhs)
    geo%conn = c_loc(conn)

    call mf_default_params(prm)
    prm%cfl = opt%cfl
    prm%vl_const = opt%vl_const
    prm%power = opt%power
    prm%mach = opt%mach
    prm%aoa = opt%aoa
    prm%euler = opt%euler
    prm%rks = opt%rks
    prm%inner_iterations = opt%inner_iterations
    prm%shapes = opt%shapes
    prm%interior_points_normal_flag = opt%interior_points_normal_flag
    prm%format = opt%format
    prm%precision = opt%precision
    prm%first_order = opt%first_order

    allocate(prim(4, pts%n))
    itr = 0
    res_old = 0.0d0
    if (opt%restart == 1) then
        write(*, *) '%%%%%%%%%%%%-Using restart file-%%%%%%%%%%%'
        write(*, *)
        if (opt%nml == 1) then
            call read_restart_dat('restart.dat', pts%n, prim, itr, res_old)
        else
            call read_restart('restart/sol.dat', pts%n, prim, itr, res_old)
        end if
        prm%restart = 1
        prm%res_old = res_old
    end if
    rc = mf_create(geo, prm, ctx)
    call check(rc, 'mf_create')
    if (opt%restart == 1) call check(mf_set_prim(ctx, prim), 'mf_set_prim')
    write(*, *) '%%%%%%%%%%%-Solution initialised-%%%%%%%%%%'
    write(*, *)
    if (opt%autotune == 1) then
        call check(mf_autotune(ctx, 10_c_int32_t, tun), 'mf_autotune')
        write(*, '(a,3i3)') ' lanes per point (lsq, inner, flux):', tun%lpp_lsq, tun%lpp_inner, tun%lpp_flux
    end if

    allocate(cl(opt%shapes), cd(opt%shapes), cm(opt%shapes), cp(max(nshape, 1)))
    call system_clock(c1)
    if (opt%ad_mode == 1) then
        ! src_mpi_adjoint: primal sweep with checkpoints, then the reverse sweep
        write(*, *) '%%%%%%%%%-Adjoint with binomial checkpoints-%%%%%%%%%%%'
        allocate(dJdx(pts%n), dJdy(pts%n))
        call check(mf_adjoint(ctx, int(itr + 1, c_int32_t), int(opt%max_iters, c_int32_t), &
                              int(opt%chkpts, c_int32_t), J, dJdx, dJdy, ast), 'mf_adjoint')
        it = itr + opt%max_iters + 1
        write(*, '(a,es24.16)') ' Objective (sum of Cd):', J
        write(*, '(a,i0,a,i0,a,i0)') ' forward steps ', ast%forward_steps, ', taped ', ast%taped_steps, &
            ', snapshots ', ast%snapshots
        call write_sensitivity('sensitivity/sensitivity.dat', pts, dJdx, dJdy, it - 1)
    else if (opt%ad_mode == 2) then
        ! src_mpi_forward (q_lskum_d.F90:28-60): pointd%x = 0 except 1 at the point with
        ! original id 1; every iteration carries the tangents; vector_cost_funcd = dCd
        write(*, *) '%%%%%%%%%%%%%-<<X Value Set To 1>>-%%%%%%%%%%%%'
        allocate(dJdx(pts%n), dJdy(pts%n))
        dJdx = 0.0d0
        dJdy = 0.0d0
        dJdx(1) = 1.0d0
        call check(mf_tangent(ctx, int(opt%max_iters, c_int32_t), dJdx, dJdy, J, Jd, c_null_ptr), 'mf_tangent')
        it = itr + opt%max_iters + 1
        write(*, '(a,es24.16)') ' Objective (sum of Cd):', J
        write(*, '(a,es24.16)') ' Vector function derivative is ', Jd
    else
        ! q_lskum.F90:59-73: per-iteration residue line and history file
        open(unit=301, file='residue', form='formatted', status='replace', action='write')
        maxchunk = max(1, min(opt%nsave, 1000))
        allocate(res(maxchunk), rnew(maxchunk), cdh(maxchunk))
        it = itr + 1
        last = itr + opt%max_iters
        do while (it <= last)
            ! chunks end on multiples of nsave, where the reference saves the solution
            chunk = min(maxchunk, last - it + 1, opt%nsave - mod(it - 1, opt%nsave))
            call check(mf_run(ctx, int(it, c_int32_t), int(chunk, c_int32_t), res, rnew, cdh, done), 'mf_run')
            do k = 1, done
                ! fpi_solver.F90:63-66: res_old follows res_new at every t <= 2
                if (it + k - 1 <= 2 .and. opt%restart == 0) res_old = rnew(k)
                write(*, '(a12,i8,a15,e30.20)') 'iterations:', it + k - 1, 'residue:', res(k)
                write(301, *) it + k - 1, res(k)
            end do
            it = it + done
            if (done < chunk) exit          ! NaN residue: q_lskum.F90:71
            if (mod(it - 1, opt%nsave) == 0) then
                write(*, *)
                write(*, *) '%%%%%%%%%%%%%-Saving solution-%%%%%%%%%%%%%'
                write(*, *)
                call check(mf_get_prim(ctx, prim), 'mf_get_prim')
                if (opt%nml == 1) then
                    call write_output_dat('output.dat', pts, prim, it - 1, res_old)
                else
                    call write_solution('solution/sol.dat', pts, prim, it - 2, res_old)
                end if
            end if
        end do
        close(301)
    end if
    call system_clock(c2)
    runtime = real(c2 - c1, 8) / real(rate, 8)

    call check(mf_compute_cl_cd_cm(ctx, cl, cd, cm, cp), 'mf_compute_cl_cd_cm')
    call write_cp('cp/cp-file', pts, cp)
    call check(mf_get_prim(ctx, prim), 'mf_get_prim')
    if (opt%nml == 1) then
        ! q_lskum.cuf:150 and meshfree_solver.F90:107: the last iteration run
        call write_output_dat('output.dat', pts, prim, it - 1, res_old)
    else
        ! header "it - 1" of the last iteration run, as fpi_solver.F90:73-82 writes it
        call write_solution('solution/sol.dat', pts, prim, it - 2, res_old)
    end if
    write(*, *) 'Vector function is ', cd
    write(*, '(a,3es24.16)') ' Cl Cd Cm (shape 1):', cl(1), cd(1), cm(1)
    call check(mf_destroy(ctx), 'mf_destroy')
    call system_clock(c2)
    totaltime = real(c2 - c0, 8) / real(rate, 8)
    write(*, *)
    write(*, *) '%%%%%%%%%%%-Simulation finished-%%%%%%%%%%%'
    write(*, *) 'Run time:  ', runtime, 'seconds'
    write(*, *) 'Total time:', totaltime, 'seconds'

contains

    subroutine check(code, what)
        integer(c_int), intent(in) :: code
        character(len=*), intent(in) :: what
        if (code /= MF_OK) then
            write(*, '(a,i0,a)') ' *** '//what//' failed (', code, '): '//mf_error_message()
            error stop 2
        end if
    end subroutine

end program meshfree_host
