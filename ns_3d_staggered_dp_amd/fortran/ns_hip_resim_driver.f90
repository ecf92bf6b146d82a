! ns_hip_resim_driver -- a Fortran host in the shape of the AB2-CN time loop
! of re_simulation_sub_tstep_interBC_combine.f90:508-802, running through
! ns_hip_mod's per-subroutine wrappers.  Each step takes the parent's
! sub-box data (u*, v*, w*, dp and the previous p, raw FP64 files written by
! the caller), extracts the wall planes with get_velpr_bc, advances the
! predictor on the host from cal_conv / cal_diff, applies the CN boundary
! rows and ADI sweeps axis by axis, forms the Poisson right-hand side, adds
! the boundary terms, solves with DCT_poisson_solver, fills the pressure
! ghosts and projects.  The calls of get_velpr_bc, get_pr_bc,
! vel_bc_staggered_CN2, pr_bc_staggered_modify_rhs, DCT_poisson_solver and
! pr_bc_staggered are written as the reference's call sites write them
! (…interBC_combine.f90:513-517,620-695,717-719,773-776), interior plane
! sections included; the dttrsb loops become adi_sweep, the inline
! divergence and projection the divergence / project wrappers.  The MKL
! trig-transform handles are placeholders: DCT_poisson_solver accepts and
! ignores them.
!
! Usage: ns_hip_resim_driver <nx> <ny> <nz> <nsteps> <dx> <dy> <dz> <nu> <dt> <dir> [mismatch]
!   <dir>/init_{u,v,w,p,hx,hy,hz}.bin            restart state (ghosted shapes)
!   <dir>/sub_{ustar,vstar,wstar,dp,ppre}_<m>.bin parent sub-box data of step m
! writes <dir>/final_{u,v,w,p}.bin, mean_dp.bin (the gauge of every step),
! velpr_1.bin (the 24 planes get_velpr_bc gave at step 1, in order) and
! pr_1.bin (get_pr_bc's 6 planes of step 1).  With "mismatch" it passes
! pr_bc_staggered a plane of the wrong shape, which must stop the run.
program ns_hip_resim_driver
    use ns_hip_mod
    implicit none
    integer :: nx, ny, nz, nsteps, nxp, nyp, nzp, t_step, ia, ib, ja, jb, ka, kb
    integer, parameter :: bc_x = 2, bc_y = 2, bc_z = 2, pbc_x = 3, pbc_y = 3, pbc_z = 3, dp_flag = 1
    logical, parameter :: using_Ustar = .true.
    real(8) :: dx, dy, dz, dx2, dy2, dz2, nu, dt, dpf
    character(len=512) :: arg, dir
    logical :: mismatch
    real(8), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:), dp(:,:,:)
    real(8), allocatable :: rhs_x_previous(:,:,:), rhs_y_previous(:,:,:), rhs_z_previous(:,:,:)
    real(8), allocatable :: u_star(:,:,:), v_star(:,:,:), w_star(:,:,:)
    real(8), allocatable :: rhs_x(:,:,:), rhs_y(:,:,:), rhs_z(:,:,:), dpdx(:,:,:), dpdy(:,:,:), dpdz(:,:,:)
    real(8), allocatable :: conv_x(:,:,:), conv_y(:,:,:), conv_z(:,:,:), diff_x(:,:,:), diff_y(:,:,:), diff_z(:,:,:)
    real(8), allocatable :: u_star_sub(:,:,:), v_star_sub(:,:,:), w_star_sub(:,:,:), dp_sub(:,:,:), p_sub_pre(:,:,:)
    real(8), allocatable :: RHS_poisson_internal(:,:,:), rhs_tt(:,:,:), eig_tt(:,:,:), mean_dps(:)
    real(8), allocatable :: bx_u_1(:,:), bx_u_nx(:,:), by_u_1(:,:), by_u_ny(:,:), bz_u_1(:,:), bz_u_nz(:,:)
    real(8), allocatable :: bx_v_1(:,:), bx_v_nx(:,:), by_v_1(:,:), by_v_ny(:,:), bz_v_1(:,:), bz_v_nz(:,:)
    real(8), allocatable :: bx_w_1(:,:), bx_w_nx(:,:), by_w_1(:,:), by_w_ny(:,:), bz_w_1(:,:), bz_w_nz(:,:)
    real(8), allocatable :: bx_p_1(:,:), bx_p_nx(:,:), by_p_1(:,:), by_p_ny(:,:), bz_p_1(:,:), bz_p_nz(:,:)
    real(8), allocatable :: qx_p_1(:,:), qx_p_nx(:,:), qy_p_1(:,:), qy_p_ny(:,:), qz_p_1(:,:), qz_p_nz(:,:)
    ! placeholders for the MKL trig-transform state the reference passes along
    integer :: handle_x = 0, handle_y = 0, handle_z = 0
    integer :: ipar_x(128), ipar_y(128), ipar_z(128)
    real(8), allocatable :: dpar_x(:), dpar_y(:), dpar_z(:)

    call get_command_argument(1, arg); read(arg, *) nx
    call get_command_argument(2, arg); read(arg, *) ny
    call get_command_argument(3, arg); read(arg, *) nz
    call get_command_argument(4, arg); read(arg, *) nsteps
    call get_command_argument(5, arg); read(arg, *) dx
    call get_command_argument(6, arg); read(arg, *) dy
    call get_command_argument(7, arg); read(arg, *) dz
    call get_command_argument(8, arg); read(arg, *) nu
    call get_command_argument(9, arg); read(arg, *) dt
    call get_command_argument(10, dir)
    call get_command_argument(11, arg)
    mismatch = trim(arg) == "mismatch"
    nxp = nx + 2; nyp = ny + 2; nzp = nz + 2
    dx2 = dx * dx; dy2 = dy * dy; dz2 = dz * dz
    dpf = real(dp_flag, 8)

    allocate(u(nx+1,ny+2,nz+2), v(nx+2,ny+1,nz+2), w(nx+2,ny+2,nz+1), p(nxp,nyp,nzp), dp(nxp,nyp,nzp))
    allocate(rhs_x_previous, mold=u); allocate(rhs_y_previous, mold=v); allocate(rhs_z_previous, mold=w)
    allocate(u_star, mold=u); allocate(v_star, mold=v); allocate(w_star, mold=w)
    allocate(rhs_x, mold=u); allocate(rhs_y, mold=v); allocate(rhs_z, mold=w)
    allocate(dpdx, mold=u); allocate(dpdy, mold=v); allocate(dpdz, mold=w)
    allocate(u_star_sub, mold=u); allocate(v_star_sub, mold=v); allocate(w_star_sub, mold=w)
    allocate(dp_sub, mold=p); allocate(p_sub_pre, mold=p)
    ! cal_conv / cal_diff outputs on the wall codes' interior ranges
    allocate(conv_x(nx-1,ny,nz), conv_y(nx,ny-1,nz), conv_z(nx,ny,nz-1))
    allocate(diff_x, mold=conv_x); allocate(diff_y, mold=conv_y); allocate(diff_z, mold=conv_z)
    allocate(RHS_poisson_internal(nx,ny,nz), rhs_tt(nx+1,ny+1,nz+1), eig_tt(nx,ny,nz), mean_dps(nsteps))
    allocate(dpar_x(5*nx/2+2), dpar_y(5*ny/2+2), dpar_z(5*nz/2+2))
    allocate(bx_u_1(ny+2,nz+2), bx_u_nx(ny+2,nz+2), by_u_1(nx+1,nz+2), by_u_ny(nx+1,nz+2))
    allocate(bz_u_1(nx+1,ny+2), bz_u_nz(nx+1,ny+2))
    allocate(bx_v_1(ny+1,nz+2), bx_v_nx(ny+1,nz+2), by_v_1(nx+2,nz+2), by_v_ny(nx+2,nz+2))
    allocate(bz_v_1(nx+2,ny+1), bz_v_nz(nx+2,ny+1))
    allocate(bx_w_1(ny+2,nz+1), bx_w_nx(ny+2,nz+1), by_w_1(nx+2,nz+1), by_w_ny(nx+2,nz+1))
    allocate(bz_w_1(nx+2,ny+2), bz_w_nz(nx+2,ny+2))
    allocate(bx_p_1(nyp,nzp), bx_p_nx(nyp,nzp), by_p_1(nxp,nzp), by_p_ny(nxp,nzp), bz_p_1(nxp,nyp), bz_p_nz(nxp,nyp))
    allocate(qx_p_1, mold=bx_p_1); allocate(qx_p_nx, mold=bx_p_1); allocate(qy_p_1, mold=by_p_1)
    allocate(qy_p_ny, mold=by_p_1); allocate(qz_p_1, mold=bz_p_1); allocate(qz_p_nz, mold=bz_p_1)
    eig_tt = 0; ipar_x = 0; ipar_y = 0; ipar_z = 0; dpar_x = 0; dpar_y = 0; dpar_z = 0
    rhs_tt = 0; dp = 0

    call load(fname("init_u"), u); call load(fname("init_v"), v); call load(fname("init_w"), w)
    call load(fname("init_p"), p)
    call load(fname("init_hx"), rhs_x_previous); call load(fname("init_hy"), rhs_y_previous)
    call load(fname("init_hz"), rhs_z_previous)

    call ns_hip_init(nx, ny, nz, dx, dy, dz, nu, dt, "AB2-CN", bc_x, bc_y, bc_z, pbc_x, pbc_y, pbc_z, dp_flag)

    ! interior ranges of the predictor (x_range0/1 etc., …interBC_combine.f90:62,161-165)
    ia = 2; ib = nx; ja = 2; jb = ny; ka = 2; kb = nz

    do t_step = 1, nsteps
        call load(fname_m("sub_ustar", t_step), u_star_sub); call load(fname_m("sub_vstar", t_step), v_star_sub)
        call load(fname_m("sub_wstar", t_step), w_star_sub); call load(fname_m("sub_dp", t_step), dp_sub)
        call load(fname_m("sub_ppre", t_step), p_sub_pre)

        call get_velpr_bc(u_star_sub, v_star_sub, w_star_sub, dp_sub, bc_x, bc_y, bc_z, pbc_x, pbc_y, pbc_z, nx, ny, nz, dx, dy, dz, &
            bx_u_1, bx_u_nx, by_u_1, by_u_ny, bz_u_1, bz_u_nz, bx_v_1, bx_v_nx, by_v_1, by_v_ny, bz_v_1, bz_v_nz, &
            bx_w_1, bx_w_nx, by_w_1, by_w_ny, bz_w_1, bz_w_nz, bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz, dt, dp_flag, using_Ustar, p_sub_pre)
        if (t_step == 1) then
            call get_pr_bc(dp_sub, pbc_x, pbc_y, pbc_z, nx, ny, nz, dx, dy, dz, qx_p_1, qx_p_nx, qy_p_1, qy_p_ny, qz_p_1, qz_p_nz)
            call dump_planes(fname("velpr_1"))
            call dump_pr(fname("pr_1"))
        end if

        call cal_conv(u, v, w, bc_x, bc_y, bc_z, dx, dy, dz, conv_x, conv_y, conv_z)
        call cal_diff(u, v, w, bc_x, bc_y, bc_z, dx2, dy2, dz2, diff_x, diff_y, diff_z)
        ! pressure gradient on the whole face arrays (…interBC_combine.f90:533-537)
        dpdx = (p(2:nxp,:,:) - p(1:nxp-1,:,:)) / dx
        dpdy = (p(:,2:nyp,:) - p(:,1:nyp-1,:)) / dy
        dpdz = (p(:,:,2:nzp) - p(:,:,1:nzp-1)) / dz

        ! AB2-CN predictor after the first step (…interBC_combine.f90:636-653):
        ! explicit half of the diffusion, AB2 convection; the history keeps -conv
        rhs_x = 0; rhs_y = 0; rhs_z = 0
        rhs_x(ia:ib, 2:ny+1, 2:nz+1) = 0.5d0 * nu * diff_x - 1.5d0 * conv_x
        rhs_y(2:nx+1, ja:jb, 2:nz+1) = 0.5d0 * nu * diff_y - 1.5d0 * conv_y
        rhs_z(2:nx+1, 2:ny+1, ka:kb) = 0.5d0 * nu * diff_z - 1.5d0 * conv_z
        u_star = dt * ((rhs_x - 0.5d0 * rhs_x_previous) - dpf * dpdx) + u
        v_star = dt * ((rhs_y - 0.5d0 * rhs_y_previous) - dpf * dpdy) + v
        w_star = dt * ((rhs_z - 0.5d0 * rhs_z_previous) - dpf * dpdz) + w
        rhs_x_previous(ia:ib, 2:ny+1, 2:nz+1) = -conv_x
        rhs_y_previous(2:nx+1, ja:jb, 2:nz+1) = -conv_y
        rhs_z_previous(2:nx+1, 2:ny+1, ka:kb) = -conv_z

        call vel_bc_staggered_CN2(u_star, v_star, w_star, bx_u_1, bx_u_nx, bx_v_1, bx_v_nx, bx_w_1, bx_w_nx, bc_x, 1)
        call adi_sweep(u_star, v_star, w_star, 1)
        call vel_bc_staggered_CN2(u_star, v_star, w_star, by_u_1, by_u_ny, by_v_1, by_v_ny, by_w_1, by_w_ny, bc_y, 2)
        call adi_sweep(u_star, v_star, w_star, 2)
        call vel_bc_staggered_CN2(u_star, v_star, w_star, bz_u_1, bz_u_nz, bz_v_1, bz_v_nz, bz_w_1, bz_w_nz, bc_z, 3)
        call adi_sweep(u_star, v_star, w_star, 3)

        call divergence(u_star, v_star, w_star, dt, RHS_poisson_internal)

        call pr_bc_staggered_modify_rhs(RHS_poisson_internal, bx_p_1(2:nyp-1,2:nzp-1), bx_p_nx(2:nyp-1,2:nzp-1), &
            by_p_1(2:nxp-1,2:nzp-1), by_p_ny(2:nxp-1,2:nzp-1), bz_p_1(2:nxp-1,2:nyp-1), bz_p_nz(2:nxp-1,2:nyp-1), &
            pbc_x, pbc_y, pbc_z, dx, dy, dz, dx2, dy2, dz2)

        rhs_tt(1:nx,1:ny,1:nz)=RHS_poisson_internal

        mean_dps(t_step) = mean(dp_sub(2:nxp-1,2:nyp-1,2:nzp-1))
        call DCT_poisson_solver(rhs_tt, eig_tt, handle_x, handle_y, handle_z, &
            ipar_x, ipar_y, ipar_z, dpar_x, dpar_y, dpar_z, nx, ny, nz, pbc_x, pbc_y, pbc_z, mean(dp_sub(2:nxp-1,2:nyp-1,2:nzp-1)))

        dp(2:nxp-1,2:nyp-1,2:nzp-1)=rhs_tt(1:nx,1:ny,1:nz)
        if (mismatch) then
            call pr_bc_staggered(dp, bx_p_1(2:nyp-1,:), bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz, pbc_x, pbc_y, pbc_z, dx, dy, dz)
        end if
        call pr_bc_staggered(dp, bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz, pbc_x, pbc_y, pbc_z, dx, dy, dz)

        call project(u_star, v_star, w_star, dp, dt, dp_flag, u, v, w, p)
    end do

    call dump(fname("final_u"), u); call dump(fname("final_v"), v); call dump(fname("final_w"), w)
    call dump(fname("final_p"), p)
    call dump1(fname("mean_dp"), mean_dps)
    print '("ns_hip_resim_driver: ", I0, "x", I0, "x", I0, ", ", I0, " AB2-CN steps")', nx, ny, nz, nsteps
    call ns_hip_finalize()

contains

    ! mean over an array section (other_utility.f90's mean)
    real(8) function mean(a)
        real(8), intent(in) :: a(:,:,:)
        mean = sum(a) / size(a)
    end function mean

    function fname(stem) result(f)
        character(*), intent(in) :: stem
        character(len=600) :: f
        f = trim(dir) // "/" // stem // ".bin"
    end function fname

    function fname_m(stem, m) result(f)
        character(*), intent(in) :: stem
        integer, intent(in) :: m
        character(len=600) :: f
        character(len=16) :: s
        write(s, '(I0)') m
        f = trim(dir) // "/" // stem // "_" // trim(s) // ".bin"
    end function fname_m

    subroutine load(f, a)
        character(*), intent(in) :: f
        real(8), intent(out) :: a(:,:,:)
        integer :: un
        open(newunit=un, file=trim(f), access='stream', form='unformatted', status='old')
        read(un) a
        close(un)
    end subroutine load

    subroutine dump(f, a)
        character(*), intent(in) :: f
        real(8), intent(in) :: a(:,:,:)
        integer :: un
        open(newunit=un, file=trim(f), access='stream', form='unformatted', status='replace')
        write(un) a
        close(un)
    end subroutine dump

    subroutine dump1(f, a)
        character(*), intent(in) :: f
        real(8), intent(in) :: a(:)
        integer :: un
        open(newunit=un, file=trim(f), access='stream', form='unformatted', status='replace')
        write(un) a
        close(un)
    end subroutine dump1

    subroutine dump_planes(f)
        character(*), intent(in) :: f
        integer :: un
        open(newunit=un, file=trim(f), access='stream', form='unformatted', status='replace')
        write(un) bx_u_1, bx_u_nx, by_u_1, by_u_ny, bz_u_1, bz_u_nz, bx_v_1, bx_v_nx, by_v_1, by_v_ny, bz_v_1, bz_v_nz
        write(un) bx_w_1, bx_w_nx, by_w_1, by_w_ny, bz_w_1, bz_w_nz, bx_p_1, bx_p_nx, by_p_1, by_p_ny, bz_p_1, bz_p_nz
        close(un)
    end subroutine dump_planes

    subroutine dump_pr(f)
        character(*), intent(in) :: f
        integer :: un
        open(newunit=un, file=trim(f), access='stream', form='unformatted', status='replace')
        write(un) qx_p_1, qx_p_nx, qy_p_1, qy_p_ny, qz_p_1, qz_p_nz
        close(un)
    end subroutine dump_pr

end program ns_hip_resim_driver
