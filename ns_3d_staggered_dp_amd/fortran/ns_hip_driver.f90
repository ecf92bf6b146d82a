! ns_hip_driver -- a Fortran host in the shape of big_periodic_3D.f90 that
! runs through ns_hip_mod: it builds a Taylor-Green field on the staggered
! grid (coordinates of big_periodic_3D.f90:14-18), calls the cal_conv /
! cal_diff mirrors, then advances the fused AB2 + FFT step and writes raw
! FP64 streams (column-major, the reference's shapes) for comparison.
! Usage: ns_hip_driver <n> <nsteps> <outdir>
program ns_hip_driver
    use ns_hip_mod
    implicit none
    real(8), parameter :: pi = 3.1415926535897932_8
    integer :: n, nsteps, i, j, k
    character(len=512) :: arg, outdir
    real(8) :: dx, nu, dt0
    real(8), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:), hx(:,:,:), hy(:,:,:), hz(:,:,:)
    real(8), allocatable :: conv_x(:,:,:), conv_y(:,:,:), conv_z(:,:,:), diff_x(:,:,:), diff_y(:,:,:), diff_z(:,:,:)

    call get_command_argument(1, arg); read(arg, *) n
    call get_command_argument(2, arg); read(arg, *) nsteps
    call get_command_argument(3, outdir)
    dx = 2.0d0 * pi / n
    nu = 1.0d0 / 1600.0d0
    dt0 = 1.0d-3
    allocate(u(n+1,n+2,n+2), v(n+2,n+1,n+2), w(n+2,n+2,n+1), p(n+2,n+2,n+2))
    allocate(hx(n+1,n+2,n+2), hy(n+2,n+1,n+2), hz(n+2,n+2,n+1))
    allocate(conv_x(n,n,n), conv_y(n,n,n), conv_z(n,n,n), diff_x(n,n,n), diff_y(n,n,n), diff_z(n,n,n))
    u = 0; v = 0; w = 0; p = 0; hx = 0; hy = 0; hz = 0
    ! u = sin x cos y cos z at (xu, yu, zu); v = -cos x sin y cos z at (xv, yv, zv)
    do k = 1, n + 2
        do j = 1, n + 2
            do i = 1, n + 1
                u(i, j, k) = sin((i - 1) * dx) * cos((j - 1.5d0) * dx) * cos((k - 1.5d0) * dx)
            end do
        end do
    end do
    do k = 1, n + 2
        do j = 1, n + 1
            do i = 1, n + 2
                v(i, j, k) = -cos((i - 1.5d0) * dx) * sin((j - 1) * dx) * cos((k - 1.5d0) * dx)
            end do
        end do
    end do
    call ns_hip_init(n, n, n, dx, dx, dx, nu, dt0, "AB2", 1, 1, 1, 1, 1, 1, 0, NS_FLAG_GAUGE_CORNER)
    call cal_conv(u, v, w, 1, 1, 1, dx, dx, dx, conv_x, conv_y, conv_z)
    call cal_diff(u, v, w, 1, 1, 1, dx * dx, dx * dx, dx * dx, diff_x, diff_y, diff_z)
    call dump(trim(outdir) // "/init_u.bin", u); call dump(trim(outdir) // "/init_v.bin", v)
    call dump(trim(outdir) // "/init_w.bin", w)
    call dump(trim(outdir) // "/conv_x.bin", conv_x); call dump(trim(outdir) // "/conv_y.bin", conv_y)
    call dump(trim(outdir) // "/diff_z.bin", diff_z)
    call ns_hip_upload(u, v, w, p, hx, hy, hz, .true.)
    call ns_hip_step(nsteps)
    call ns_hip_download(u, v, w, p, hx, hy, hz)
    call dump(trim(outdir) // "/final_u.bin", u); call dump(trim(outdir) // "/final_v.bin", v)
    call dump(trim(outdir) // "/final_w.bin", w); call dump(trim(outdir) // "/final_p.bin", p)
    print '("ns_hip_driver: n=", I0, " steps=", I0, " max div ", ES12.4)', n, nsteps, ns_hip_max_div()
    call ns_hip_finalize()
contains
    subroutine dump(fname, a)
        character(*), intent(in) :: fname
        real(8), intent(in) :: a(:,:,:)
        integer :: un
        open(newunit=un, file=fname, access='stream', form='unformatted', status='replace')
        write(un) a
        close(un)
    end subroutine dump
end program ns_hip_driver
