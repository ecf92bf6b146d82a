! ns_hip_mod -- ISO_C_BINDING layer over libns_hip (include/ns_hip.h).
!
! A Fortran host of NS_3D_Staggered_dp replaces `use NS_functions` by
! `use ns_hip_mod` and keeps its calls: the wrappers below carry the
! reference's subroutine names and argument lists (NS_functions.f90):
!   cal_conv            :53-59      cal_diff              :161-168
!   vel_bc_staggered    :213-217    vel_bc_staggered_CN2  :313
!   get_velpr_bc        :414-416    get_pr_bc             :514
!   pr_bc_staggered_modify_rhs :1345-1347   pr_bc_staggered :1403-1405
!   DCT_poisson_solver  :1505-1507 (eig_tt, the MKL handles and ipar/dpar
!                                   are accepted and not used: the
!                                   eigenvalues and transform plans live in
!                                   the context; (rhs, mean_dp) also works)
! plus the inline driver pieces as subroutines (divergence, poisson_fft,
! project, adi_sweep for the dttrsb/getrs loops) and the fused,
! device-resident step (ns_hip_step).  Sizes, codes and spacings passed
! explicitly must equal the context's (ns_hip_init).
! All arrays keep the reference's ghosted shapes.  Errors stop with the
! failing entry point's name and status, as the drivers' `stop(1)` does.
module ns_hip_mod
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer, parameter, public :: NS_EULER = 0, NS_AB2 = 1, NS_AB2_CN = 2
    integer, parameter, public :: NS_FLAG_GAUGE_CORNER = 1, NS_FLAG_CN_CYCLIC = 2, NS_FLAG_NO_GRAPH = 4, &
        NS_FLAG_EAGER_PROJECTION = 8

    type, bind(C), public :: ns_config
        integer(c_int) :: nx, ny, nz
        real(c_double) :: dx, dy, dz
        real(c_double) :: nu, dt
        integer(c_int) :: scheme
        integer(c_int) :: bc(3), pbc(3)
        integer(c_int) :: dp_flag
        integer(c_int) :: flags
    end type ns_config

    interface
        integer(c_int) function c_ns_create(cfg, ctx) bind(C, name='ns_create')
            import :: c_int, c_ptr, ns_config
            type(ns_config), intent(in) :: cfg
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function c_ns_destroy(ctx) bind(C, name='ns_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_long) function c_ns_plane_set_length(nx, ny, nz) bind(C, name='ns_plane_set_length')
            import :: c_int, c_long
            integer(c_int), value :: nx, ny, nz
        end function
        integer(c_int) function c_ns_upload_state(ctx, u, v, w, p, hx, hy, hz, first) bind(C, name='ns_upload_state')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, p, hx, hy, hz
            integer(c_int), value :: first
        end function
        integer(c_int) function c_ns_download_state(ctx, u, v, w, p, hx, hy, hz) bind(C, name='ns_download_state')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, p, hx, hy, hz
        end function
        integer(c_int) function c_ns_download_work(ctx, us, vs, ws, dp) bind(C, name='ns_download_work')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, us, vs, ws, dp
        end function
        integer(c_int) function c_ns_set_bc_planes(ctx, planes) bind(C, name='ns_set_bc_planes')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, planes
        end function
        integer(c_int) function c_ns_step(ctx, nsteps) bind(C, name='ns_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function c_ns_max_div(ctx, out) bind(C, name='ns_max_div')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out
        end function
        integer(c_int) function c_ns_set_dt(ctx, dt) bind(C, name='ns_set_dt')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function c_ns_history_save(ctx) bind(C, name='ns_history_save')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_ns_history_restore(ctx) bind(C, name='ns_history_restore')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_ns_bc_interp_spline(ctx, snaps, nsnap, snap_dt, t) bind(C, name='ns_bc_interp_spline')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, snaps
            integer(c_int), value :: nsnap
            real(c_double), value :: snap_dt, t
        end function
        integer(c_int) function c_ns_error_metrics(ctx, u, v, w, p, err, rms) bind(C, name='ns_error_metrics')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, p
            real(c_double), intent(out) :: err(8), rms(4)
        end function
        integer(c_int) function c_ns_cal_conv(ctx, u, v, w, bx, by, bz, dx, dy, dz, cx, cy, cz) bind(C, name='ns_cal_conv')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, cx, cy, cz
            integer(c_int), value :: bx, by, bz
            real(c_double), value :: dx, dy, dz
        end function
        integer(c_int) function c_ns_cal_diff(ctx, u, v, w, bx, by, bz, dx2, dy2, dz2, fx, fy, fz) bind(C, name='ns_cal_diff')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, fx, fy, fz
            integer(c_int), value :: bx, by, bz
            real(c_double), value :: dx2, dy2, dz2
        end function
        integer(c_int) function c_ns_vel_bc(ctx, u, v, w, planes, bx, by, bz) bind(C, name='ns_vel_bc_staggered')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, planes
            integer(c_int), value :: bx, by, bz
        end function
        integer(c_int) function c_ns_vel_bc_cn2(ctx, u, v, w, planes, bc, dir) bind(C, name='ns_vel_bc_staggered_cn2')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w, planes
            integer(c_int), value :: bc, dir
        end function
        integer(c_int) function c_ns_adi_sweep(ctx, u, v, w, dir) bind(C, name='ns_adi_sweep')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, u, v, w
            integer(c_int), value :: dir
        end function
        integer(c_int) function c_ns_get_velpr_bc(ctx, u, v, w, p, dt, dp_flag, using_ustar, pprev, planes) &
            bind(C, name='ns_get_velpr_bc')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, p, pprev, planes
            real(c_double), value :: dt
            integer(c_int), value :: dp_flag, using_ustar
        end function
        integer(c_int) function c_ns_get_pr_bc(ctx, p, planes) bind(C, name='ns_get_pr_bc')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, p, planes
        end function
        integer(c_int) function c_ns_divergence(ctx, u, v, w, dt, rhs) bind(C, name='ns_divergence')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, u, v, w, rhs
            real(c_double), value :: dt
        end function
        integer(c_int) function c_ns_modify_rhs(ctx, rhs, planes) bind(C, name='ns_pr_bc_staggered_modify_rhs')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, rhs, planes
        end function
        integer(c_int) function c_ns_poisson_fft(ctx, rhs, sol) bind(C, name='ns_poisson_fft')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, rhs, sol
        end function
        integer(c_int) function c_ns_dct_poisson(ctx, rhs, mean_dp) bind(C, name='ns_dct_poisson_solver')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, rhs
            real(c_double), value :: mean_dp
        end function
        integer(c_int) function c_ns_pr_bc(ctx, dp, planes) bind(C, name='ns_pr_bc_staggered')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, dp, planes
        end function
        integer(c_int) function c_ns_project(ctx, us, vs, ws, dp, dt, dp_flag, u, v, w, p) bind(C, name='ns_project')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, us, vs, ws, dp, u, v, w, p
            real(c_double), value :: dt
            integer(c_int), value :: dp_flag
        end function
    end interface

    type(c_ptr), save :: ctx = c_null_ptr
    type(ns_config), save :: cfg_saved
    real(c_double), allocatable, target, save :: pset(:)

    public :: ns_hip_init, ns_hip_finalize, ns_hip_upload, ns_hip_download, ns_hip_download_work, ns_hip_step
    public :: ns_hip_set_planes, ns_hip_max_div, ns_hip_set_dt, ns_hip_history_save, ns_hip_history_restore
    public :: ns_hip_bc_interp_spline, ns_hip_error_metrics
    public :: cal_conv, cal_diff, vel_bc_staggered, vel_bc_staggered_CN2, get_velpr_bc, get_pr_bc
    public :: pr_bc_staggered_modify_rhs, pr_bc_staggered, DCT_poisson_solver, poisson_fft, divergence, project
    public :: adi_sweep

    ! the reference's argument list, or (rhs, mean_dp)
    interface DCT_poisson_solver
        module procedure dct_poisson_solver_ref, dct_poisson_solver_ctx
    end interface DCT_poisson_solver

contains

    subroutine check(rc, name)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: name
        if (rc /= 0) then
            print '("ns_hip: ", A, " failed with status ", I0)', name, rc
            stop 1
        end if
    end subroutine check

    ! timescheme: "Euler" / "AB2" / "AB2-CN" (big_periodic_3D.f90:105)
    subroutine ns_hip_init(nx, ny, nz, dx, dy, dz, nu, dt, timescheme, bc_x, bc_y, bc_z, pbc_x, pbc_y, pbc_z, dp_flag, flags)
        integer, intent(in) :: nx, ny, nz, bc_x, bc_y, bc_z, pbc_x, pbc_y, pbc_z, dp_flag
        real(8), intent(in) :: dx, dy, dz, nu, dt
        character(*), intent(in) :: timescheme
        integer, intent(in), optional :: flags
        type(ns_config) :: c
        if (c_associated(ctx)) call ns_hip_finalize()
        c%nx = nx; c%ny = ny; c%nz = nz
        c%dx = dx; c%dy = dy; c%dz = dz; c%nu = nu; c%dt = dt
        select case (trim(timescheme))
        case ("Euler"); c%scheme = NS_EULER
        case ("AB2"); c%scheme = NS_AB2
        case ("AB2-CN"); c%scheme = NS_AB2_CN
        case default
            print *, "ns_hip_init: unknown time scheme ", trim(timescheme)
            stop 1
        end select
        c%bc = [bc_x, bc_y, bc_z]
        c%pbc = [pbc_x, pbc_y, pbc_z]
        c%dp_flag = dp_flag
        c%flags = 0
        if (present(flags)) c%flags = flags
        call check(c_ns_create(c, ctx), "ns_create")
        cfg_saved = c
        if (allocated(pset)) deallocate(pset)
        allocate(pset(c_ns_plane_set_length(nx, ny, nz)))
        pset = 0
    end subroutine ns_hip_init

    subroutine ns_hip_finalize()
        if (c_associated(ctx)) call check(c_ns_destroy(ctx), "ns_destroy")
        ctx = c_null_ptr
    end subroutine ns_hip_finalize

    subroutine ns_hip_upload(u, v, w, p, rhs_x_previous, rhs_y_previous, rhs_z_previous, first)
        real(8), intent(in), target, contiguous :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:)
        real(8), intent(in), target, contiguous :: rhs_x_previous(:,:,:), rhs_y_previous(:,:,:), rhs_z_previous(:,:,:)
        logical, intent(in) :: first
        integer(c_int) :: f
        f = 0
        if (first) f = 1
        call check(c_ns_upload_state(ctx, c_loc(u), c_loc(v), c_loc(w), c_loc(p), c_loc(rhs_x_previous), &
            c_loc(rhs_y_previous), c_loc(rhs_z_previous), f), "ns_upload_state")
    end subroutine ns_hip_upload

    subroutine ns_hip_download(u, v, w, p, rhs_x_previous, rhs_y_previous, rhs_z_previous)
        real(8), intent(inout), target, contiguous :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:)
        real(8), intent(inout), target, contiguous :: rhs_x_previous(:,:,:), rhs_y_previous(:,:,:), rhs_z_previous(:,:,:)
        call check(c_ns_download_state(ctx, c_loc(u), c_loc(v), c_loc(w), c_loc(p), c_loc(rhs_x_previous), &
            c_loc(rhs_y_previous), c_loc(rhs_z_previous)), "ns_download_state")
    end subroutine ns_hip_download

    subroutine ns_hip_download_work(u_star, v_star, w_star, dp)
        real(8), intent(inout), target, contiguous :: u_star(:,:,:), v_star(:,:,:), w_star(:,:,:), dp(:,:,:)
        call check(c_ns_download_work(ctx, c_loc(u_star), c_loc(v_star), c_loc(w_star), c_loc(dp)), "ns_download_work")
    end subroutine ns_hip_download_work

    subroutine ns_hip_set_planes(planes)
        real(8), intent(in), target, contiguous :: planes(:)
        call check(c_ns_set_bc_planes(ctx, c_loc(planes)), "ns_set_bc_planes")
    end subroutine ns_hip_set_planes

    subroutine ns_hip_step(nsteps)
        integer, intent(in) :: nsteps
        call check(c_ns_step(ctx, int(nsteps, c_int)), "ns_step")
    end subroutine ns_hip_step

    real(8) function ns_hip_max_div()
        real(c_double) :: m
        call check(c_ns_max_div(ctx, m), "ns_max_div")
        ns_hip_max_div = m
    end function ns_hip_max_div

    ! sub-step phase of re_simulation_sub_tstep_interBC_combine.f90:446-448,602-612
    subroutine ns_hip_set_dt(dt)
        real(8), intent(in) :: dt
        call check(c_ns_set_dt(ctx, dt), "ns_set_dt")
    end subroutine ns_hip_set_dt

    subroutine ns_hip_history_save()
        call check(c_ns_history_save(ctx), "ns_history_save")
    end subroutine ns_hip_history_save

    subroutine ns_hip_history_restore()
        call check(c_ns_history_restore(ctx), "ns_history_restore")
    end subroutine ns_hip_history_restore

    ! plane set := interp1(..., 'spline') of the snapshot history planes(:, 1:nsnap) at time t
    subroutine ns_hip_bc_interp_spline(snaps, nsnap, snap_dt, t)
        real(8), intent(in), target, contiguous :: snaps(:,:)
        integer, intent(in) :: nsnap
        real(8), intent(in) :: snap_dt, t
        call check(c_ns_bc_interp_spline(ctx, c_loc(snaps), int(nsnap, c_int), snap_dt, t), "ns_bc_interp_spline")
    end subroutine ns_hip_bc_interp_spline

    ! err_vel(1:8, t_step) and rms_vel(1:4, t_step) of …interBC_combine.f90:826-833
    subroutine ns_hip_error_metrics(u_sub, v_sub, w_sub, p_sub, err_vel, rms_vel)
        real(8), intent(in), target, contiguous :: u_sub(:,:,:), v_sub(:,:,:), w_sub(:,:,:), p_sub(:,:,:)
        real(8), intent(out) :: err_vel(8), rms_vel(4)
        call check(c_ns_error_metrics(ctx, c_loc(u_sub), c_loc(v_sub), c_loc(w_sub), c_loc(p_sub), err_vel, rms_vel), &
            "ns_error_metrics")
    end subroutine ns_hip_error_metrics

    ! ---- mirrors with the reference's argument lists ----
    !
    ! Grid sizes, codes and spacings passed explicitly (as the reference's
    ! routines take them) must equal the context's: the kernels use the
    ! context's.  Arrays passed to the library are `contiguous` dummies, so an
    ! array section is copied in (and out) by the compiler; planes are checked
    ! against their shapes in the plane set.

    subroutine fail(name, what)
        character(*), intent(in) :: name, what
        print '("ns_hip: ", A, ": ", A)', name, what
        stop 1
    end subroutine fail

    subroutine expect_grid(name, nx, ny, nz)
        character(*), intent(in) :: name
        integer, intent(in) :: nx, ny, nz
        if (nx /= cfg_saved%nx .or. ny /= cfg_saved%ny .or. nz /= cfg_saved%nz) &
            call fail(name, "nx, ny, nz differ from the context's")
    end subroutine expect_grid

    subroutine expect_spacing(name, dx, dy, dz)
        character(*), intent(in) :: name
        real(8), intent(in) :: dx, dy, dz
        if (dx /= cfg_saved%dx .or. dy /= cfg_saved%dy .or. dz /= cfg_saved%dz) &
            call fail(name, "dx, dy, dz differ from the context's")
    end subroutine expect_spacing

    subroutine expect_codes(name, kind, cx, cy, cz)
        character(*), intent(in) :: name
        integer, intent(in) :: kind, cx, cy, cz  ! kind 1: velocity codes, 2: pressure codes
        integer :: c(3)
        c = cfg_saved%bc
        if (kind == 2) c = cfg_saved%pbc
        if (cx /= c(1) .or. cy /= c(2) .or. cz /= c(3)) call fail(name, "boundary codes differ from the context's")
    end subroutine expect_codes

    subroutine expect_shape3(name, a, n1, n2, n3)
        character(*), intent(in) :: name
        real(8), intent(in) :: a(:,:,:)
        integer, intent(in) :: n1, n2, n3
        if (size(a, 1) /= n1 .or. size(a, 2) /= n2 .or. size(a, 3) /= n3) call fail(name, "array shape mismatch")
    end subroutine expect_shape3

    subroutine expect_velocity(name, u, v, w)
        character(*), intent(in) :: name
        real(8), intent(in) :: u(:,:,:), v(:,:,:), w(:,:,:)
        integer :: nx, ny, nz
        nx = cfg_saved%nx; ny = cfg_saved%ny; nz = cfg_saved%nz
        call expect_shape3(name, u, nx + 1, ny + 2, nz + 2)
        call expect_shape3(name, v, nx + 2, ny + 1, nz + 2)
        call expect_shape3(name, w, nx + 2, ny + 2, nz + 1)
    end subroutine expect_velocity

    subroutine cal_conv(u, v, w, bc_x, bc_y, bc_z, dx, dy, dz, conv_x, conv_y, conv_z)
        integer, intent(in) :: bc_x, bc_y, bc_z
        real(8), intent(in) :: dx, dy, dz
        real(8), dimension(:,:,:), intent(in), target, contiguous :: u, v, w
        real(8), dimension(:,:,:), intent(out), target, contiguous :: conv_x, conv_y, conv_z
        call expect_velocity("cal_conv", u, v, w)
        call check(c_ns_cal_conv(ctx, c_loc(u), c_loc(v), c_loc(w), int(bc_x, c_int), int(bc_y, c_int), &
            int(bc_z, c_int), dx, dy, dz, c_loc(conv_x), c_loc(conv_y), c_loc(conv_z)), "cal_conv")
    end subroutine cal_conv

    subroutine cal_diff(u, v, w, bc_x, bc_y, bc_z, dx2, dy2, dz2, diff_x, diff_y, diff_z)
        integer, intent(in) :: bc_x, bc_y, bc_z
        real(8), intent(in) :: dx2, dy2, dz2
        real(8), dimension(:,:,:), intent(in), target, contiguous :: u, v, w
        real(8), dimension(:,:,:), intent(out), target, contiguous :: diff_x, diff_y, diff_z
        call expect_velocity("cal_diff", u, v, w)
        call check(c_ns_cal_diff(ctx, c_loc(u), c_loc(v), c_loc(w), int(bc_x, c_int), int(bc_y, c_int), &
            int(bc_z, c_int), dx2, dy2, dz2, c_loc(diff_x), c_loc(diff_y), c_loc(diff_z)), "cal_diff")
    end subroutine cal_diff

    ! dimensions of plane q (0-based order of include/ns_hip.h: u, v, w, p;
    ! per component x, y, z; per axis the low then the high end)
    subroutine plane_dims(q, m1, m2)
        integer, intent(in) :: q
        integer, intent(out) :: m1, m2
        integer :: d(3), c
        c = q / 6
        d = [cfg_saved%nx + 2, cfg_saved%ny + 2, cfg_saved%nz + 2]
        if (c < 3) d(c + 1) = d(c + 1) - 1
        select case (mod(q, 6) / 2)
        case (0); m1 = d(2); m2 = d(3)
        case (1); m1 = d(1); m2 = d(3)
        case default; m1 = d(1); m2 = d(2)
        end select
    end subroutine plane_dims

    integer function plane_offset(q)
        integer, intent(in) :: q
        integer :: k, m1, m2
        plane_offset = 0
        do k = 0, q - 1
            call plane_dims(k, m1, m2)
            plane_offset = plane_offset + m1 * m2
        end do
    end function plane_offset

    ! the whole plane q := b (its shape must be the plane's)
    subroutine put_plane(q, b, name)
        integer, intent(in) :: q
        real(8), intent(in) :: b(:,:)
        character(*), intent(in) :: name
        integer :: o, m1, m2
        call plane_dims(q, m1, m2)
        if (size(b, 1) /= m1 .or. size(b, 2) /= m2) call fail(name, "plane shape mismatch")
        o = plane_offset(q)
        pset(o + 1:o + m1 * m2) = reshape(b, [m1 * m2])
    end subroutine put_plane

    ! plane q := b given either whole or as its interior section (2:m1-1, 2:m2-1),
    ! the form the reference passes to pr_bc_staggered_modify_rhs
    subroutine put_plane_or_interior(q, b, name)
        integer, intent(in) :: q
        real(8), intent(in) :: b(:,:)
        character(*), intent(in) :: name
        integer :: o, m1, m2, k
        call plane_dims(q, m1, m2)
        if (size(b, 1) == m1 .and. size(b, 2) == m2) then
            call put_plane(q, b, name)
        else if (size(b, 1) == m1 - 2 .and. size(b, 2) == m2 - 2) then
            o = plane_offset(q)
            do k = 2, m2 - 1
                pset(o + (k - 1) * m1 + 2:o + (k - 1) * m1 + m1 - 1) = b(:, k - 1)
            end do
        else
            call fail(name, "plane shape mismatch (neither the whole plane nor its interior)")
        end if
    end subroutine put_plane_or_interior

    subroutine get_plane(q, b, name)
        integer, intent(in) :: q
        real(8), intent(out) :: b(:,:)
        character(*), intent(in) :: name
        integer :: o, m1, m2
        call plane_dims(q, m1, m2)
        if (size(b, 1) /= m1 .or. size(b, 2) /= m2) call fail(name, "plane shape mismatch")
        o = plane_offset(q)
        b = reshape(pset(o + 1:o + m1 * m2), [m1, m2])
    end subroutine get_plane

    subroutine vel_bc_staggered(u_star, v_star, w_star, &
        bx_u_1, bx_u_nx, by_u_1, by_u_ny, bz_u_1, bz_u_nz, &
        bx_v_1, bx_v_nx, by_v_1, by_v_ny, bz_v_1, bz_v_nz, &
        bx_w_1, bx_w_nx, by_w_1, by_w_ny, bz_w_1, bz_w_nz, bc_x, bc_y, bc_z)
        integer, intent(in) :: bc_x, bc_y, bc_z
        real(8), dimension(:,:,:), intent(inout), target, contiguous :: u_star, v_star, w_star
        real(8), dimension(:,:), intent(in) :: bx_u_1, bx_u_nx, by_u_1, by_u_ny, bz_u_1, bz_u_nz
        real(8), dimension(:,:), intent(in) :: bx_v_1, bx_v_nx, by_v_1, by_v_ny, bz_v_1, bz_v_nz
        real(8), dimension(:,:), intent(in) :: bx_w_1, bx_w_nx, by_w_1, by_w_ny, bz_w_1, bz_w_nz
        character(*), parameter :: nm = "vel_bc_staggered"
        call expect_velocity(nm, u_star, v_star, w_star)
        call put_plane(0, bx_u_1, nm); call put_plane(1, bx_u_nx, nm); call put_plane(2, by_u_1, nm)
        call put_plane(3, by_u_ny, nm); call put_plane(4, bz_u_1, nm); call put_plane(5, bz_u_nz, nm)
        call put_plane(6, bx_v_1, nm); call put_plane(7, bx_v_nx, nm); call put_plane(8, by_v_1, nm)
        call put_plane(9, by_v_ny, nm); call put_plane(10, bz_v_1, nm); call put_plane(11, bz_v_nz, nm)
        call put_plane(12, bx_w_1, nm); call put_plane(13, bx_w_nx, nm); call put_plane(14, by_w_1, nm)
        call put_plane(15, by_w_ny, nm); call put_plane(16, bz_w_1, nm); call put_plane(17, bz_w_nz, nm)
        call check(c_ns_vel_bc(ctx, c_loc(u_star), c_loc(v_star), c_loc(w_star), c_loc(pset), int(bc_x, c_int), &
            int(bc_y, c_int), int(bc_z, c_int)), nm)
    end subroutine vel_bc_staggered

    subroutine vel_bc_staggered_CN2(u_star, v_star, w_star, u_1, u_end, v_1, v_end, w_1, w_end, bc, dir)
        integer, intent(in) :: bc, dir
        real(8), dimension(:,:,:), intent(inout), target, contiguous :: u_star, v_star, w_star
        real(8), dimension(:,:), intent(in) :: u_1, u_end, v_1, v_end, w_1, w_end
        character(*), parameter :: nm = "vel_bc_staggered_CN2"
        integer :: a
        if (dir < 1 .or. dir > 3) call fail(nm, "dir must be 1, 2 or 3")
        call expect_velocity(nm, u_star, v_star, w_star)
        a = 2 * (dir - 1)
        if (bc == 2 .or. bc == 4) then
            call put_plane(a, u_1, nm); call put_plane(a + 1, u_end, nm)
            call put_plane(6 + a, v_1, nm); call put_plane(7 + a, v_end, nm)
            call put_plane(12 + a, w_1, nm); call put_plane(13 + a, w_end, nm)
        end if
        call check(c_ns_vel_bc_cn2(ctx, c_loc(u_star), c_loc(v_star), c_loc(w_star), c_loc(pset), int(bc, c_int), &
            int(dir, c_int)), nm)
    end subroutine vel_bc_staggered_CN2

    ! the ADI sweep along dir with the context's CN line systems: the
    ! dttrsb/getrs loops of …interBC_combine.f90:656-701 / big_periodic_3D.f90:425-483
    subroutine adi_sweep(u_star, v_star, w_star, dir)
        integer, intent(in) :: dir
        real(8), dimension(:,:,:), intent(inout), target, contiguous :: u_star, v_star, w_star
        call expect_velocity("adi_sweep", u_star, v_star, w_star)
        call check(c_ns_adi_sweep(ctx, c_loc(u_star), c_loc(v_star), c_loc(w_star), int(dir, c_int)), "adi_sweep")
    end subroutine adi_sweep

    ! NS_functions.f90:414-416
    subroutine get_velpr_bc(u_bc, v_bc, w_bc, p_bc, bc_x, bc_y, bc_z, pbc_x, pbc_y, pbc_z, nx, ny, nz, dx, dy, dz, &
        bx_u_1, bx_u_nx, by_u_1, by_u_ny, bz_u_1, bz_u_nz, bx_v_1, bx_v_nx, by_v_1, by_v_ny, bz_v_1, bz_v_nz, &
        bx_w_1, bx_w_nx, by_w_1, by_w_ny, bz_w_1, bz_w_nz, bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz, &
        dt, dp_flag, using_Ustar, p_previous)
        integer, intent(in) :: bc_x, bc_y, bc_z, pbc_x, pbc_y, pbc_z, nx, ny, nz, dp_flag
        real(8), intent(in) :: dx, dy, dz, dt
        real(8), intent(in), target, contiguous :: u_bc(:,:,:), v_bc(:,:,:), w_bc(:,:,:), p_bc(:,:,:), &
            p_previous(:,:,:)
        real(8), dimension(:,:), intent(out) :: bx_u_1, bx_u_nx, by_u_1, by_u_ny, bz_u_1, bz_u_nz
        real(8), dimension(:,:), intent(out) :: bx_v_1, bx_v_nx, by_v_1, by_v_ny, bz_v_1, bz_v_nz
        real(8), dimension(:,:), intent(out) :: bx_w_1, bx_w_nx, by_w_1, by_w_ny, bz_w_1, bz_w_nz
        real(8), dimension(:,:), intent(out) :: bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz
        logical, intent(in) :: using_Ustar
        character(*), parameter :: nm = "get_velpr_bc"
        integer(c_int) :: us
        call expect_grid(nm, nx, ny, nz)
        call expect_spacing(nm, dx, dy, dz)
        call expect_codes(nm, 1, bc_x, bc_y, bc_z)
        call expect_codes(nm, 2, pbc_x, pbc_y, pbc_z)
        call expect_velocity(nm, u_bc, v_bc, w_bc)
        call expect_shape3(nm, p_bc, nx + 2, ny + 2, nz + 2)
        call expect_shape3(nm, p_previous, nx + 2, ny + 2, nz + 2)
        us = 0
        if (using_Ustar) us = 1
        pset = 0
        call check(c_ns_get_velpr_bc(ctx, c_loc(u_bc), c_loc(v_bc), c_loc(w_bc), c_loc(p_bc), dt, int(dp_flag, c_int), &
            us, c_loc(p_previous), c_loc(pset)), nm)
        call get_plane(0, bx_u_1, nm); call get_plane(1, bx_u_nx, nm); call get_plane(2, by_u_1, nm)
        call get_plane(3, by_u_ny, nm); call get_plane(4, bz_u_1, nm); call get_plane(5, bz_u_nz, nm)
        call get_plane(6, bx_v_1, nm); call get_plane(7, bx_v_nx, nm); call get_plane(8, by_v_1, nm)
        call get_plane(9, by_v_ny, nm); call get_plane(10, bz_v_1, nm); call get_plane(11, bz_v_nz, nm)
        call get_plane(12, bx_w_1, nm); call get_plane(13, bx_w_nx, nm); call get_plane(14, by_w_1, nm)
        call get_plane(15, by_w_ny, nm); call get_plane(16, bz_w_1, nm); call get_plane(17, bz_w_nz, nm)
        call get_plane(18, bx_p_1, nm); call get_plane(19, bx_p_nx, nm); call get_plane(20, by_p_1, nm)
        call get_plane(21, by_p_ny, nm); call get_plane(22, bz_p_1, nm); call get_plane(23, bz_p_nz, nm)
    end subroutine get_velpr_bc

    ! NS_functions.f90:514-521
    subroutine get_pr_bc(p_bc, pbc_x, pbc_y, pbc_z, nx, ny, nz, dx, dy, dz, bx_p_1, bx_p_nx, by_p_1, by_p_ny, &
        bz_p_1, bz_p_nz)
        integer, intent(in) :: pbc_x, pbc_y, pbc_z, nx, ny, nz
        real(8), intent(in) :: dx, dy, dz
        real(8), intent(in), target, contiguous :: p_bc(:,:,:)
        real(8), dimension(:,:), intent(out) :: bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz
        character(*), parameter :: nm = "get_pr_bc"
        call expect_grid(nm, nx, ny, nz)
        call expect_spacing(nm, dx, dy, dz)
        call expect_codes(nm, 2, pbc_x, pbc_y, pbc_z)
        call expect_shape3(nm, p_bc, nx + 2, ny + 2, nz + 2)
        pset = 0
        call check(c_ns_get_pr_bc(ctx, c_loc(p_bc), c_loc(pset)), nm)
        call get_plane(18, bx_p_1, nm); call get_plane(19, bx_p_nx, nm); call get_plane(20, by_p_1, nm)
        call get_plane(21, by_p_ny, nm); call get_plane(22, bz_p_1, nm); call get_plane(23, bz_p_nz, nm)
    end subroutine get_pr_bc

    ! NS_functions.f90:1345-1347.  The planes may be whole or, as the
    ! reference's call site passes them (…interBC_combine.f90:717-719),
    ! their interior sections (2:nyp-1, 2:nzp-1) etc.
    subroutine pr_bc_staggered_modify_rhs(RHS_poisson_internal, bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz, &
        pbc_x, pbc_y, pbc_z, dx, dy, dz, dx2, dy2, dz2)
        integer, intent(in) :: pbc_x, pbc_y, pbc_z
        real(8), intent(inout), target, contiguous :: RHS_poisson_internal(:,:,:)
        real(8), dimension(:,:), intent(in) :: bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz
        real(8), intent(in) :: dx, dy, dz, dx2, dy2, dz2
        character(*), parameter :: nm = "pr_bc_staggered_modify_rhs"
        call expect_spacing(nm, dx, dy, dz)
        if (dx2 /= dx * dx .or. dy2 /= dy * dy .or. dz2 /= dz * dz) call fail(nm, "dx2, dy2, dz2 are not dx*dx, ...")
        call expect_codes(nm, 2, pbc_x, pbc_y, pbc_z)
        call expect_shape3(nm, RHS_poisson_internal, cfg_saved%nx, cfg_saved%ny, cfg_saved%nz)
        call put_plane_or_interior(18, bx_p_1, nm); call put_plane_or_interior(19, bx_p_nx, nm)
        call put_plane_or_interior(20, by_p_1, nm); call put_plane_or_interior(21, by_p_ny, nm)
        call put_plane_or_interior(22, bz_p_1, nm); call put_plane_or_interior(23, bz_p_nz, nm)
        call check(c_ns_modify_rhs(ctx, c_loc(RHS_poisson_internal), c_loc(pset)), nm)
    end subroutine pr_bc_staggered_modify_rhs

    ! NS_functions.f90:1403-1405
    subroutine pr_bc_staggered(dp, bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz, pbc_x, pbc_y, pbc_z, dx, dy, dz)
        integer, intent(in) :: pbc_x, pbc_y, pbc_z
        real(8), intent(inout), target, contiguous :: dp(:,:,:)
        real(8), dimension(:,:), intent(in) :: bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz
        real(8), intent(in) :: dx, dy, dz
        character(*), parameter :: nm = "pr_bc_staggered"
        call expect_spacing(nm, dx, dy, dz)
        call expect_codes(nm, 2, pbc_x, pbc_y, pbc_z)
        call expect_shape3(nm, dp, cfg_saved%nx + 2, cfg_saved%ny + 2, cfg_saved%nz + 2)
        call put_plane(18, bx_p_1, nm); call put_plane(19, bx_p_nx, nm); call put_plane(20, by_p_1, nm)
        call put_plane(21, by_p_ny, nm); call put_plane(22, bz_p_1, nm); call put_plane(23, bz_p_nz, nm)
        call check(c_ns_pr_bc(ctx, c_loc(dp), c_loc(pset)), nm)
    end subroutine pr_bc_staggered

    ! DCT_poisson_solver(rhs, mean_dp): rhs(nx,ny,nz) in place
    subroutine dct_poisson_solver_ctx(rhs, mean_dp)
        real(8), intent(inout), target, contiguous :: rhs(:,:,:)
        real(8), intent(in) :: mean_dp
        call expect_shape3("DCT_poisson_solver", rhs, cfg_saved%nx, cfg_saved%ny, cfg_saved%nz)
        call check(c_ns_dct_poisson(ctx, c_loc(rhs), mean_dp), "DCT_poisson_solver")
    end subroutine dct_poisson_solver_ctx

    ! NS_functions.f90:1505-1507, the reference's argument list: rhs_tt is
    ! (nx+1,ny+1,nz+1) with the data in (1:nx,1:ny,1:nz) (or (nx,ny,nz));
    ! eig_tt, the MKL trig-transform handles and ipar/dpar are accepted and
    ! not used (the context holds its own eigenvalues and transform plans), so
    ! the call site compiles unchanged
    subroutine dct_poisson_solver_ref(rhs_tt, eig_tt, handle_x, handle_y, handle_z, ipar_x, ipar_y, ipar_z, &
        dpar_x, dpar_y, dpar_z, nx, ny, nz, pbc_x, pbc_y, pbc_z, mean_dp)
        integer, intent(in) :: nx, ny, nz, pbc_x, pbc_y, pbc_z
        real(8), intent(inout) :: rhs_tt(:,:,:)
        ! assumed rank: any shape matches (the reference's are rank 3 and 1)
        real(8), intent(in) :: eig_tt(..), dpar_x(..), dpar_y(..), dpar_z(..)
        type(*), intent(in) :: handle_x, handle_y, handle_z
        integer, intent(in) :: ipar_x(..), ipar_y(..), ipar_z(..)
        real(8), intent(in) :: mean_dp
        character(*), parameter :: nm = "DCT_poisson_solver"
        real(8), allocatable, target :: r(:,:,:)
        call expect_grid(nm, nx, ny, nz)
        call expect_codes(nm, 2, pbc_x, pbc_y, pbc_z)
        if (size(rhs_tt, 1) < nx .or. size(rhs_tt, 2) < ny .or. size(rhs_tt, 3) < nz) call fail(nm, "rhs_tt too small")
        allocate(r(nx, ny, nz))
        r = rhs_tt(1:nx, 1:ny, 1:nz)
        call check(c_ns_dct_poisson(ctx, c_loc(r), mean_dp), nm)
        rhs_tt(1:nx, 1:ny, 1:nz) = r
    end subroutine dct_poisson_solver_ref

    ! periodic FFT Poisson solve (big_periodic_3D.f90:526-548): sol = lap^-1 rhs, mean zero
    subroutine poisson_fft(rhs, sol)
        real(8), intent(in), target, contiguous :: rhs(:,:,:)
        real(8), intent(out), target, contiguous :: sol(:,:,:)
        call expect_shape3("poisson_fft", rhs, cfg_saved%nx, cfg_saved%ny, cfg_saved%nz)
        call expect_shape3("poisson_fft", sol, cfg_saved%nx, cfg_saved%ny, cfg_saved%nz)
        call check(c_ns_poisson_fft(ctx, c_loc(rhs), c_loc(sol)), "poisson_fft")
    end subroutine poisson_fft

    ! RHS_poisson_internal = div(u*)/dt (…interBC_combine.f90:709-712)
    subroutine divergence(u_star, v_star, w_star, dt, rhs)
        real(8), intent(in), target, contiguous :: u_star(:,:,:), v_star(:,:,:), w_star(:,:,:)
        real(8), intent(in) :: dt
        real(8), intent(out), target, contiguous :: rhs(:,:,:)
        call expect_velocity("divergence", u_star, v_star, w_star)
        call expect_shape3("divergence", rhs, cfg_saved%nx, cfg_saved%ny, cfg_saved%nz)
        call check(c_ns_divergence(ctx, c_loc(u_star), c_loc(v_star), c_loc(w_star), dt, c_loc(rhs)), "divergence")
    end subroutine divergence

    ! u = u* - dt*diff(dp)/dx ..., p = dp_flag*p + dp (…interBC_combine.f90:795-801)
    subroutine project(u_star, v_star, w_star, dp, dt, dp_flag, u, v, w, p)
        real(8), intent(in), target, contiguous :: u_star(:,:,:), v_star(:,:,:), w_star(:,:,:), dp(:,:,:)
        real(8), intent(in) :: dt
        integer, intent(in) :: dp_flag
        real(8), intent(inout), target, contiguous :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:)
        call expect_velocity("project", u_star, v_star, w_star)
        call expect_velocity("project", u, v, w)
        call check(c_ns_project(ctx, c_loc(u_star), c_loc(v_star), c_loc(w_star), c_loc(dp), dt, int(dp_flag, c_int), &
            c_loc(u), c_loc(v), c_loc(w), c_loc(p)), "project")
    end subroutine project

end module ns_hip_mod
