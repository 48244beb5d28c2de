! Host-side I/O of the meshfree solver, in the reference's formats.
!
!  * case.in: PETSc-style options, one "-key value" per line, with readcase.F90's
!    names and defaults (src_mpi_primal/readcase.F90:10-125; the adjoint's
!    -adjoint_mode/-chkpts, src_mpi_adjoint/readcase.F90:16-50).
!  * point/point.h5 (point_preprocessor.F90:215-546) through the HDF5 C library
!    (the HDF5 Fortran modules of this image were built by another Fortran compiler),
!    else the legacy text point/partGrid (:23-146).
!  * restart/sol.dat (initial_conditions.F90:44-69).
!  * residue (q_lskum.F90:59-73), cp/cp-file (compute_force_coeffs.F90:29-74),
!    solution/sol.dat (post_processing.F90:11-38) and sensitivity/sensitivity.dat
!    (src_mpi_adjoint/post_processing.F90:41-66).
module host_io
    use iso_c_binding
    implicit none

    type case_options
        real(8) :: cfl = 0.0d0, mach = 0.63d0, aoa = 2.0d0, power = 0.0d0, vl_const = 150.0d0
        real(8) :: tfinal = 1.0d20
        integer :: max_iters = 10000000, inner_iterations = 0, interior_points_normal_flag = 0
        integer :: shapes = 1, nsave = 1000000, restart = 0, format = 1, rks = 4
        real(8) :: euler = 1.0d0
        integer :: first_order = 0, ad_mode = 0, chkpts = 0
        ! options with no counterpart in readcase: the device path's precision and
        ! whether to time the kernels' launch shapes before running
        integer :: precision = 0, autotune = 1
        ! 1: options from input.nml (the CUDA build's run: output.dat, restart.dat)
        integer :: nml = 0
        character(len=64) :: time = 'steady', tscheme = 'ssprk43', solution_accuracy = 'second'
        character(len=64) :: format_file = 'legacy', restart_solution = 'no', adjoint_mode = 'blackbox'
        character(len=64) :: limiter = 'venkat', run_option = 'normal', precision_name = 'fp64'
    end type

    type point_data
        integer :: n = 0, w = 0, ghost = 0
        integer, allocatable :: original_id(:), left(:), right(:), flag_1(:), flag_2(:), nbhs(:), qtdepth(:)
        integer, allocatable :: conn(:, :)          ! conn(n, w), 1-based, as the reference stores it
        real(8), allocatable :: x(:), y(:), nx(:), ny(:), min_dist(:)
    end type

    ! HDF5 C API (hid_t = int64_t, herr_t = int, hsize_t = uint64)
    integer(c_int64_t), parameter :: H5P_DEFAULT = 0, H5S_ALL = 0
    integer(c_int), parameter :: H5F_ACC_RDONLY = 0, H5T_INTEGER = 0, H5T_FLOAT = 1
    interface
        integer(c_int) function h5open() bind(C, name="H5open")
            import
        end function
        integer(c_int64_t) function h5fopen(name, flags, fapl) bind(C, name="H5Fopen")
            import
            character(kind=c_char), intent(in) :: name(*)
            integer(c_int), value :: flags
            integer(c_int64_t), value :: fapl
        end function
        integer(c_int) function h5fclose(id) bind(C, name="H5Fclose")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int64_t) function h5gopen2(loc, name, gapl) bind(C, name="H5Gopen2")
            import
            integer(c_int64_t), value :: loc, gapl
            character(kind=c_char), intent(in) :: name(*)
        end function
        integer(c_int) function h5gclose(id) bind(C, name="H5Gclose")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int64_t) function h5aopen(obj, name, aapl) bind(C, name="H5Aopen")
            import
            integer(c_int64_t), value :: obj, aapl
            character(kind=c_char), intent(in) :: name(*)
        end function
        integer(c_int64_t) function h5aget_type(id) bind(C, name="H5Aget_type")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int) function h5aread(id, mtype, buf) bind(C, name="H5Aread")
            import
            integer(c_int64_t), value :: id, mtype
            type(c_ptr), value :: buf
        end function
        integer(c_int) function h5aclose(id) bind(C, name="H5Aclose")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int64_t) function h5dopen2(loc, name, dapl) bind(C, name="H5Dopen2")
            import
            integer(c_int64_t), value :: loc, dapl
            character(kind=c_char), intent(in) :: name(*)
        end function
        integer(c_int64_t) function h5dget_type(id) bind(C, name="H5Dget_type")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int64_t) function h5dget_space(id) bind(C, name="H5Dget_space")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int) function h5dread(id, mtype, mspace, fspace, xfer, buf) bind(C, name="H5Dread")
            import
            integer(c_int64_t), value :: id, mtype, mspace, fspace, xfer
            type(c_ptr), value :: buf
        end function
        integer(c_int) function h5dclose(id) bind(C, name="H5Dclose")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int) function h5sget_simple_extent_dims(id, dims, maxdims) bind(C, name="H5Sget_simple_extent_dims")
            import
            integer(c_int64_t), value :: id
            integer(c_int64_t), intent(out) :: dims(*)
            type(c_ptr), value :: maxdims
        end function
        integer(c_int) function h5sclose(id) bind(C, name="H5Sclose")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_size_t) function h5tget_size(id) bind(C, name="H5Tget_size")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int) function h5tget_class(id) bind(C, name="H5Tget_class")
            import
            integer(c_int64_t), value :: id
        end function
        integer(c_int) function h5tclose(id) bind(C, name="H5Tclose")
            import
            integer(c_int64_t), value :: id
        end function
    end interface

contains

    subroutine fail(msg)
        character(len=*), intent(in) :: msg
        write(*, '(a)') ' *** '//trim(msg)
        error stop 1
    end subroutine

    ! -------------------------------------------------------------------------
    ! case.in
    subroutine read_case(path, o)
        character(len=*), intent(in) :: path
        type(case_options), intent(inout) :: o
        character(len=512) :: line
        character(len=128) :: key, val
        integer :: u, ios, p
        open(newunit=u, file=path, status='old', action='read', iostat=ios)
        if (ios /= 0) call fail('cannot open '//trim(path))
        do
            read(u, '(a)', iostat=ios) line
            if (ios /= 0) exit
            line = adjustl(line)
            if (len_trim(line) == 0 .or. line(1:1) /= '-') cycle
            p = index(line, ' ')
            key = line(2:p - 1)
            val = adjustl(line(p:))
            call set_option(o, trim(key), trim(val))
        end do
        close(u)
        call finish_options(o)
    end subroutine

    subroutine set_option(o, key, val)
        type(case_options), intent(inout) :: o
        character(len=*), intent(in) :: key, val
        select case (key)
        case ('cfl');                           read(val, *) o%cfl
        case ('mach');                          read(val, *) o%mach
        case ('aoa');                           read(val, *) o%aoa
        case ('max_iters');                     read(val, *) o%max_iters
        case ('tfinal');                        read(val, *) o%tfinal
        case ('time');                          o%time = val
        case ('run_option');                    o%run_option = val
        case ('tscheme');                       o%tscheme = val
        case ('power');                         read(val, *) o%power
        case ('limiter');                       o%limiter = val
        case ('vl_const');                      read(val, *) o%vl_const
        case ('restart_solution');              o%restart_solution = val
        case ('inner_iterations');              read(val, *) o%inner_iterations
        case ('interior_points_normal_flag');   read(val, *) o%interior_points_normal_flag
        case ('shapes');                        read(val, *) o%shapes
        case ('nsave');                         read(val, *) o%nsave
        case ('solution_accuracy');             o%solution_accuracy = val
        case ('format_file');                   o%format_file = val
        case ('adjoint_mode');                  o%adjoint_mode = val
        case ('chkpts');                        read(val, *) o%chkpts
        case ('precision');                     o%precision_name = val
        case ('autotune');                      o%autotune = merge(1, 0, val == 'yes')
        case default
            write(*, '(a)') ' ignoring option -'//key
        end select
    end subroutine

    subroutine finish_options(o)
        type(case_options), intent(inout) :: o
        if (trim(o%tscheme) == 'first') then
            o%rks = 1
            o%euler = 2.0d0
        else if (trim(o%tscheme) == 'ssprk43') then
            o%rks = 4
            o%euler = 1.0d0
        else
            call fail('unknown -tscheme '//trim(o%tscheme))
        end if
        o%restart = merge(1, 0, trim(o%restart_solution) == 'same')
        o%first_order = merge(1, 0, trim(o%solution_accuracy) == 'first')
        o%format = merge(2, 1, trim(o%format_file) == 'quadtree')
        ! -adjoint_mode checkpoints -chkpts S: binomial schedule of S snapshots
        ! (q_lskum_chkpts_b); -adjoint_mode tape (no counterpart in the reference): every
        ! iteration taped during the primal sweep (MF_ADJ_TAPE_ALL), when it fits in HBM
        o%ad_mode = merge(1, 0, trim(o%adjoint_mode) == 'checkpoints' .or. trim(o%adjoint_mode) == 'tape')
        if (trim(o%adjoint_mode) == 'tape') o%chkpts = -1
        ! -adjoint_mode forward: the forward (tangent) mode of src_mpi_forward, seeded as
        ! q_lskum_d.F90:34-39 seeds it (x of the point with original id 1)
        if (trim(o%adjoint_mode) == 'forward') o%ad_mode = 2
        if (trim(o%precision_name) == 'mixed') then
            o%precision = 1
        else if (trim(o%precision_name) == 'fp64') then
            o%precision = 0
        else
            call fail('unknown -precision '//trim(o%precision_name))
        end if
        if (trim(o%time) /= 'steady') call fail('-time unsteady is not supported')
        if (trim(o%limiter) /= 'venkat') call fail('only -limiter venkat is supported')
        if (o%ad_mode == 1 .and. o%chkpts < 1 .and. trim(o%adjoint_mode) /= 'tape') &
            call fail('-adjoint_mode checkpoints needs -chkpts >= 1')
    end subroutine

    ! input.nml of the CUDA build (src_cuda/data_structure.F90:106-176): namelist
    ! input_parameters with that build's defaults (:52-104).  blockx, blocky and blockz
    ! (its thread-block shape) are read and have no effect: mf_autotune picks the
    ! kernels' launch shapes.
    subroutine read_nml(path, o)
        character(len=*), intent(in) :: path
        type(case_options), intent(inout) :: o
        integer :: shapes, max_iters, blockx, blocky, blockz, nsave, interior_points_normal_flag, inner_iterations
        real(8) :: cfl, vl_const, power, mach, aoa
        character(len=20) :: restart_solution, solution_accuracy, format_file, tscheme
        namelist /input_parameters/ shapes, cfl, max_iters, blockx, blocky, blockz, vl_const, power, &
            restart_solution, solution_accuracy, format_file, nsave, interior_points_normal_flag, tscheme, &
            mach, aoa, inner_iterations
        integer :: u, ios
        shapes = 1
        cfl = 0.0d0
        max_iters = 10000000
        blockx = 32
        blocky = 1
        blockz = 1
        vl_const = 150.0d0
        power = 0.0d0
        restart_solution = 'no'
        solution_accuracy = 'second'
        format_file = 'legacy'
        nsave = 10000000
        interior_points_normal_flag = 0
        tscheme = 'first'
        mach = 0.0d0
        aoa = 0.0d0
        inner_iterations = 0
        open(newunit=u, file=path, status='old', action='read', iostat=ios)
        if (ios /= 0) call fail('cannot open '//trim(path))
        read(u, nml=input_parameters, iostat=ios)
        if (ios /= 0) call fail('cannot read namelist input_parameters from '//trim(path))
        close(u)
        write(*, nml=input_parameters)
        o%nml = 1
        o%shapes = shapes
        o%cfl = cfl
        o%max_iters = max_iters
        o%vl_const = vl_const
        o%power = power
        o%nsave = nsave
        o%interior_points_normal_flag = interior_points_normal_flag
        o%inner_iterations = inner_iterations
        o%mach = mach
        o%aoa = aoa
        o%tscheme = tscheme
        o%solution_accuracy = solution_accuracy
        o%format_file = format_file
        ! data_structure.F90:150-154: 'yes' restarts (from restart.dat)
        o%restart_solution = merge('same', 'no  ', trim(restart_solution) == 'yes')
        call finish_options(o)
        ! data_structure.F90:168-172: quadtree files carry their own normals
        if (o%format == 2) o%interior_points_normal_flag = 100000
    end subroutine

    subroutine print_options(o)
        type(case_options), intent(in) :: o
        write(*, *) 'max_iters       :', o%max_iters
        write(*, *) 'cfl         :', o%cfl
        write(*, *) 'mach        :', o%mach
        write(*, *) 'aoa         :', o%aoa
        write(*, *) 'shapes          :', o%shapes
        write(*, *) 'timestep        :', trim(o%time)
        write(*, *) 'run_option      :', trim(o%run_option)
        write(*, *) 'tscheme         :', trim(o%tscheme)
        write(*, *) 'file format     :', trim(o%format_file)
        write(*, *) 'accuracy        :', trim(o%solution_accuracy)
        write(*, *) 'nsave           :', o%nsave
        write(*, *) 'power           :', o%power
        write(*, *) 'vl_const        :', o%vl_const
        write(*, *) 'inner_iterations    :', o%inner_iterations
        write(*, *) 'restart         :', trim(o%restart_solution)
        write(*, *) 'precision       :', trim(o%precision_name)
        if (o%ad_mode == 1) write(*, *) 'checkpoints         :', o%chkpts, '(-1: every iteration taped)'
    end subroutine

    ! -------------------------------------------------------------------------
    ! point files
    logical function file_exists(path)
        character(len=*), intent(in) :: path
        inquire(file=path, exist=file_exists)
    end function

    subroutine alloc_points(p, n, w)
        type(point_data), intent(inout) :: p
        integer, intent(in) :: n, w
        p%n = n
        p%w = w
        allocate(p%original_id(n), p%left(n), p%right(n), p%flag_1(n), p%flag_2(n), p%nbhs(n), p%qtdepth(n))
        allocate(p%conn(n, w), p%x(n), p%y(n), p%nx(n), p%ny(n), p%min_dist(n))
        p%conn = 0
        p%nx = 0.0d0
        p%ny = 1.0d0
        p%qtdepth = 0
    end subroutine

    subroutine read_points(o, p)
        type(case_options), intent(in) :: o
        type(point_data), intent(inout) :: p
        if (file_exists('point/point.h5')) then
            call read_point_h5('point/point.h5', p)
        else if (file_exists('point/partGrid')) then
            call read_partgrid('point/partGrid', o%format, p)
        else
            call fail('neither point/point.h5 nor point/partGrid found')
        end if
        call check_points(o, p)
    end subroutine

    subroutine check_points(o, p)
        type(case_options), intent(in) :: o
        type(point_data), intent(in) :: p
        integer :: k
        do k = 1, p%n
            if (p%flag_2(k) > o%shapes) call fail('shapes value wrong, check again')
        end do
    end subroutine

    function cstr(s) result(c)
        character(len=*), intent(in) :: s
        character(kind=c_char) :: c(len_trim(s) + 1)
        integer :: i
        do i = 1, len_trim(s)
            c(i) = s(i:i)
        end do
        c(len_trim(s) + 1) = c_null_char
    end function

    integer function read_int_attr(g, name) result(v)
        integer(c_int64_t), intent(in) :: g
        character(len=*), intent(in) :: name
        integer(c_int64_t) :: a, t
        integer(c_int32_t), target :: v4
        integer(c_int64_t), target :: v8
        a = h5aopen(g, cstr(name), H5P_DEFAULT)
        if (a < 0) call fail('point.h5: no attribute '//name)
        t = h5aget_type(a)
        if (h5tget_class(t) /= H5T_INTEGER) call fail('point.h5: attribute '//name//' is not an integer')
        if (h5tget_size(t) == 4) then
            if (h5aread(a, t, c_loc(v4)) < 0) call fail('point.h5: reading '//name)
            v = v4
        else if (h5tget_size(t) == 8) then
            if (h5aread(a, t, c_loc(v8)) < 0) call fail('point.h5: reading '//name)
            v = int(v8)
        else
            call fail('point.h5: attribute '//name//' has an unexpected size')
        end if
        if (h5tclose(t) < 0 .or. h5aclose(a) < 0) call fail('point.h5: closing '//name)
    end function

    ! read_hdf5input_point_data (point_preprocessor.F90:215-546), one partition
    subroutine read_point_h5(path, p)
        character(len=*), intent(in) :: path
        type(point_data), intent(inout) :: p
        integer(c_int64_t) :: f, g, d, sp, t
        integer(c_int64_t) :: dims(2)
        real(c_double), allocatable, target :: tab(:, :)
        integer :: n_total, n_local, n_ghost, ncol, k, r, w
        if (h5open() < 0) call fail('H5open')
        f = h5fopen(cstr(path), H5F_ACC_RDONLY, H5P_DEFAULT)
        if (f < 0) call fail('cannot open '//trim(path))
        g = h5gopen2(f, cstr('/1'), H5P_DEFAULT)
        if (g < 0) call fail(trim(path)//': no group /1')
        n_total = read_int_attr(g, 'total')
        n_local = read_int_attr(g, 'local')
        n_ghost = read_int_attr(g, 'ghost')
        if (n_total /= n_local) call fail(trim(path)//': several partitions (total /= local)')
        d = h5dopen2(g, cstr('local'), H5P_DEFAULT)
        if (d < 0) call fail(trim(path)//': no dataset /1/local')
        sp = h5dget_space(d)
        if (h5sget_simple_extent_dims(sp, dims, c_null_ptr) /= 2) call fail('/1/local is not 2-D')
        ! C order [local][ncols]: dims(1) = local, dims(2) = ncols
        ncol = int(dims(2))
        if (int(dims(1)) < n_local .or. ncol < 12) call fail('/1/local has the wrong shape')
        t = h5dget_type(d)
        if (h5tget_class(t) /= H5T_FLOAT .or. h5tget_size(t) /= 8) call fail('/1/local is not float64')
        allocate(tab(ncol, int(dims(1))))
        if (h5dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, c_loc(tab)) < 0) call fail('reading /1/local')
        if (h5tclose(t) < 0 .or. h5sclose(sp) < 0 .or. h5dclose(d) < 0) call fail('closing /1/local')
        if (h5gclose(g) < 0 .or. h5fclose(f) < 0) call fail('closing '//trim(path))
        w = 0
        do k = 1, n_local
            w = max(w, int(tab(12, k)))
        end do
        if (w > ncol - 12) call fail('/1/local: nbhs exceeds the connectivity columns')
        call alloc_points(p, n_local, max(w, 1))
        p%ghost = n_ghost
        do k = 1, n_local
            p%original_id(k) = int(tab(1, k))
            p%x(k) = tab(2, k)
            p%y(k) = tab(3, k)
            p%nx(k) = tab(4, k)
            p%ny(k) = tab(5, k)
            p%min_dist(k) = tab(6, k)
            p%left(k) = int(tab(7, k))
            p%right(k) = int(tab(8, k))
            p%qtdepth(k) = int(tab(9, k))
            p%flag_1(k) = int(tab(10, k))
            p%flag_2(k) = int(tab(11, k))
            p%nbhs(k) = int(tab(12, k))
            do r = 1, p%nbhs(k)
                p%conn(k, r) = int(tab(12 + r, k))
            end do
        end do
    end subroutine

    ! read_input_point_data (point_preprocessor.F90:23-146), one partition
    subroutine read_partgrid(path, fmt, p)
        character(len=*), intent(in) :: path
        integer, intent(in) :: fmt
        type(point_data), intent(inout) :: p
        integer :: u, ios, nproc, max_points, local_points, ghost_points, k, r, w, pass
        integer :: id, l, rr, f1, f2, nb, qd
        real(8) :: x, y, nx, ny, md
        integer :: c(4096)
        open(newunit=u, file=path, status='old', action='read', iostat=ios)
        if (ios /= 0) call fail('cannot open '//trim(path))
        ! two passes: the widest stencil first, then the data
        do pass = 1, 2
            rewind(u)
            read(u, *) nproc, max_points, local_points, ghost_points
            if (nproc /= 1) call fail(trim(path)//': several partitions')
            if (pass == 2) then
                call alloc_points(p, local_points, max(w, 1))
                p%ghost = ghost_points
            end if
            w = 0
            do k = 1, local_points
                nx = 0.0d0
                ny = 1.0d0
                qd = 0
                if (fmt == 1) then
                    read(u, *) id, x, y, l, rr, f1, f2, md, nb, (c(r), r = 1, nb)
                else
                    read(u, *) id, x, y, l, rr, f1, f2, nx, ny, qd, md, nb, (c(r), r = 1, nb)
                end if
                w = max(w, nb)
                if (pass == 2) then
                    p%original_id(k) = id
                    p%x(k) = x
                    p%y(k) = y
                    p%left(k) = l
                    p%right(k) = rr
                    p%flag_1(k) = f1
                    p%flag_2(k) = f2
                    p%nx(k) = nx
                    p%ny(k) = ny
                    p%qtdepth(k) = qd
                    p%min_dist(k) = md
                    p%nbhs(k) = nb
                    p%conn(k, 1:nb) = c(1:nb)
                end if
            end do
        end do
        close(u)
    end subroutine

    ! restart_sol (initial_conditions.F90:44-69): header (n, itr, res_old), then per
    ! point "id flag_1 flag_2 x y rho u v p"
    subroutine read_restart(path, n, prim, itr, res_old)
        character(len=*), intent(in) :: path
        integer, intent(in) :: n
        real(8), intent(out) :: prim(4, n)
        integer, intent(out) :: itr
        real(8), intent(out) :: res_old
        integer :: u, ios, i, dummy, nfile
        real(8) :: dx, dy
        open(newunit=u, file=path, status='old', action='read', iostat=ios)
        if (ios /= 0) call fail('cannot open '//trim(path))
        read(u, *) nfile, itr, res_old
        if (nfile /= n) call fail(trim(path)//': point count differs from the point file')
        do i = 1, n
            read(u, *) dummy, dummy, dummy, dx, dy, prim(1, i), prim(2, i), prim(3, i), prim(4, i)
        end do
        close(u)
    end subroutine

    ! -------------------------------------------------------------------------
    ! writers
    subroutine write_solution(path, p, prim, it_minus_1, res_old)
        character(len=*), intent(in) :: path
        type(point_data), intent(in) :: p
        real(8), intent(in) :: prim(4, p%n)
        integer, intent(in) :: it_minus_1
        real(8), intent(in) :: res_old
        integer :: u, i
        open(newunit=u, file=path)
        write(u, *) p%n, it_minus_1, res_old
        do i = 1, p%n
            write(u, '(3i8,6e30.20)') p%original_id(i), p%flag_1(i), p%flag_2(i), p%x(i), p%y(i), &
                prim(1, i), prim(2, i), prim(3, i), prim(4, i)
        end do
        close(u)
    end subroutine

    ! output.dat of the CUDA build (src_cuda/post_processing.F90:7-36): "N it res_old",
    ! then per point x, y, flag_1, qtdepth, prim, the Mach value as that routine
    ! computes it (sqrt(|u|^2 / a), a the speed of sound: kept as written there), the
    ! entropy measure (compute_entropy.F90:8-28) and the adaptation sensor
    ! (adaptation_sensors.F90:7-101).  Both are evaluated from this prim at every
    ! save; the CUDA build evaluates them once, after its last iteration.
    subroutine write_output_dat(path, p, prim, it, res_old)
        character(len=*), intent(in) :: path
        type(point_data), intent(in) :: p
        real(8), intent(in) :: prim(4, p%n)
        integer, intent(in) :: it
        real(8), intent(in) :: res_old
        real(8), allocatable :: entropy(:), sensor(:)
        real(8), parameter :: gamma = 1.4d0, pr_inf = 1.0d0 / 1.4d0
        real(8) :: sos, vel_mag, mach_number
        integer :: u, i
        allocate(entropy(p%n), sensor(p%n))
        call entropy_measure(p, prim, entropy)
        call adapt_sensor(p, prim, sensor)
        open(newunit=u, file=path)
        write(u, *) p%n, it, res_old
        do i = 1, p%n
            sos = dsqrt(gamma * prim(4, i) / prim(1, i))
            vel_mag = prim(2, i) * prim(2, i) + prim(3, i) * prim(3, i)
            mach_number = dsqrt(vel_mag / sos)
            write(u, '(2e30.20,2i4,7e30.20)') p%x(i), p%y(i), p%flag_1(i), p%qtdepth(i), &
                prim(1, i), prim(2, i), prim(3, i), prim(4, i), mach_number, entropy(i), sensor(i)
        end do
        close(u)
    end subroutine

    ! compute_entropy.F90:8-28: |ln(p / rho^gamma) - ln(p_inf)| per point
    subroutine entropy_measure(p, prim, entropy)
        type(point_data), intent(in) :: p
        real(8), intent(in) :: prim(4, p%n)
        real(8), intent(out) :: entropy(p%n)
        real(8), parameter :: gamma = 1.4d0, pr_inf = 1.0d0 / 1.4d0
        real(8) :: temp1, temp2
        integer :: k
        temp2 = dlog(pr_inf)
        do k = 1, p%n
            temp1 = prim(1, k)**gamma
            temp1 = prim(4, k) / temp1
            temp1 = dlog(temp1)
            entropy(k) = dabs(temp1 - temp2)
        end do
    end subroutine

    ! adaptation_sensors.F90:7-101: per point the largest D2 distance to a neighbour's
    ! state, normalised by the largest over all points
    subroutine adapt_sensor(p, prim, sensor)
        type(point_data), intent(in) :: p
        real(8), intent(in) :: prim(4, p%n)
        real(8), intent(out) :: sensor(p%n)
        real(8) :: u1_i, u2_i, pr_i, rho_i, u1_j, u2_j, pr_j, rho_j
        real(8) :: maxi, u_sqr, temp1, temp2, temp3, d2, max_sensor
        integer :: i, j, r
        do i = 1, p%n
            u1_i = prim(2, i)
            u2_i = prim(3, i)
            rho_i = prim(1, i)
            pr_i = prim(4, i)
            maxi = 0.0
            do r = 1, p%nbhs(i)
                j = p%conn(i, r)
                u1_j = prim(2, j)
                u2_j = prim(3, j)
                rho_j = prim(1, j)
                pr_j = prim(4, j)
                temp1 = (pr_i * pr_i * rho_j) / (pr_j * pr_j * rho_i)
                temp1 = dlog(temp1)
                temp1 = (rho_j - rho_i) * temp1
                u_sqr = (u1_j - u1_i)**2 + (u2_j - u2_i)**2
                temp2 = (rho_j / rho_i) + (pr_i * rho_j) / (pr_j * rho_i)
                temp2 = temp2 * u_sqr
                temp2 = (0.5 * rho_i * rho_i / pr_i) * temp2
                temp3 = rho_i * (((pr_i * rho_j) / (pr_j * rho_i)) - 1)
                temp3 = temp3 + rho_j * (((pr_j * rho_i) / (pr_i * rho_j)) - 1)
                temp3 = 2 * temp3
                d2 = temp1 + temp2 + temp3
                if (d2 > maxi) maxi = d2
            end do
            sensor(i) = maxi
        end do
        max_sensor = 0.0d0
        do i = 1, p%n
            if (sensor(i) >= max_sensor) max_sensor = sensor(i)
        end do
        do i = 1, p%n
            sensor(i) = sensor(i) / max_sensor
        end do
    end subroutine

    ! restart.dat of the CUDA build (initial_conditions.F90:26-37): output.dat's layout,
    ! prim in columns 5-8
    subroutine read_restart_dat(path, n, prim, itr, res_old)
        character(len=*), intent(in) :: path
        integer, intent(in) :: n
        real(8), intent(out) :: prim(4, n)
        integer, intent(out) :: itr
        real(8), intent(out) :: res_old
        integer :: u, ios, i, nfile, f1, qt
        real(8) :: dx, dy, m, e, sn
        open(newunit=u, file=path, status='old', action='read', iostat=ios)
        if (ios /= 0) call fail('cannot open '//trim(path))
        read(u, *) nfile, itr, res_old
        if (nfile /= n) call fail(trim(path)//': point count differs from the point file')
        do i = 1, n
            read(u, *) dx, dy, f1, qt, prim(1, i), prim(2, i), prim(3, i), prim(4, i), m, e, sn
        end do
        close(u)
    end subroutine

    ! cp/cp-file: "flag_2 x cp" per shape point, in point order
    subroutine write_cp(path, p, cp)
        character(len=*), intent(in) :: path
        type(point_data), intent(in) :: p
        real(8), intent(in) :: cp(*)
        integer :: u, k, j
        open(newunit=u, file=path, status='replace', action='write')
        j = 0
        do k = 1, p%n
            if (p%flag_2(k) > 0) then
                j = j + 1
                write(u, *) p%flag_2(k), p%x(k), cp(j)
            end if
        end do
        close(u)
    end subroutine

    subroutine write_sensitivity(path, p, dJdx, dJdy, it_minus_1)
        character(len=*), intent(in) :: path
        type(point_data), intent(in) :: p
        real(8), intent(in) :: dJdx(p%n), dJdy(p%n)
        integer, intent(in) :: it_minus_1
        integer :: u, i
        open(newunit=u, file=path)
        write(u, *) p%n, it_minus_1
        do i = 1, p%n
            write(u, '(1i8, 2e30.20)') p%original_id(i), dJdx(i), dJdy(i)
        end do
        close(u)
    end subroutine

end module host_io
