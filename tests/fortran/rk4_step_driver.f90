! One RK4 time step of a uniform periodic 2D forest, driven from Fortran
! through the wabbit_hip bind(C) module: the calls WABBIT's timeStep_tree
! would make on the device path (pool, heavy data in the host's
! hvy_block(nx, ny, nz, neq, Nb) layout, active list with origins/spacings,
! RHS ghost patches, calculate_time_step's block minimum, one
! RungeKuttaGeneric step, download).  The inputs are files a test writes:
!   params.bin   one wh_acm_params (its C memory image)
!   butcher.bin  the RK tableau, nrows x nrows doubles, row after row
!   hvy.bin      hvy_block(nx, ny, 1, 3, nb)
!   x0dx.bin     origin(3) and spacing(3) per active block
!   patches.bin  npatch wh_patch records
! and the outputs dt.bin (the block dt minimum) and out.bin (hvy_block after
! the step).  Arguments: directory, nb, npatch, Bs, g.
program rk4_step_driver
    use wabbit_hip
    implicit none
    character(len=1024) :: dir, arg
    integer(c_int) :: nb, npatch, bs, g, nx, i
    type(wh_pool_desc) :: d
    type(wh_acm_params) :: p
    type(wh_patch), allocatable :: pt(:)
    real(c_double), allocatable :: hvy(:, :, :, :, :), x0dx(:, :)
    integer(c_int32_t), allocatable :: act(:)
    integer(c_int8_t), allocatable :: raw(:)
    real(c_double) :: bt(25), dt
    type(c_ptr) :: pool
    integer :: u

    call get_command_argument(1, dir)
    call get_command_argument(2, arg)
    read(arg, *) nb
    call get_command_argument(3, arg)
    read(arg, *) npatch
    call get_command_argument(4, arg)
    read(arg, *) bs
    call get_command_argument(5, arg)
    read(arg, *) g
    nx = bs + 2 * g
    allocate(hvy(nx, nx, 1, 3, nb), x0dx(6, nb), act(nb), pt(npatch))

    ! the C memory image, padding included (reading the derived type itself
    ! would transfer its components one after another, without the padding)
    allocate(raw(c_sizeof(p)))
    open(newunit=u, file=trim(dir) // '/params.bin', access='stream', form='unformatted', status='old')
    read(u) raw
    close(u)
    p = transfer(raw, p)
    open(newunit=u, file=trim(dir) // '/butcher.bin', access='stream', form='unformatted', status='old')
    read(u) bt
    close(u)
    open(newunit=u, file=trim(dir) // '/hvy.bin', access='stream', form='unformatted', status='old')
    read(u) hvy
    close(u)
    open(newunit=u, file=trim(dir) // '/x0dx.bin', access='stream', form='unformatted', status='old')
    read(u) x0dx
    close(u)
    open(newunit=u, file=trim(dir) // '/patches.bin', access='stream', form='unformatted', status='old')
    read(u) pt
    close(u)
    do i = 1, nb
        act(i) = i - 1   ! hvy ids are 0-based on the library's side
    end do

    d%dim = 2
    d%bs = [bs, bs, 1]
    d%g = g
    d%neq = 3
    d%nb_max = nb
    d%n_work_slots = 3
    d%n_tmp_eqn = 0
    d%n_mask = 0
    d%mask_per_block = 1
    call wh_check(wh_pool_create(pool, d), "pool_create")
    call wh_check(wh_pool_upload(pool, WH_BLOCK, 0, nb, hvy), "pool_upload")
    call wh_check(wh_set_active(pool, act, nb, x0dx), "set_active")
    call wh_check(wh_set_patches(pool, 0, pt, npatch), "set_patches")
    call wh_check(wh_dt_reduce(pool, p, dt), "dt_reduce")
    call wh_check(wh_rk_step(pool, p, bt, 5, 0.0_c_double, dt, 2), "rk_step")
    call wh_check(wh_pool_download(pool, WH_BLOCK, 0, nb, hvy), "pool_download")
    call wh_check(wh_pool_destroy(pool), "pool_destroy")

    open(newunit=u, file=trim(dir) // '/dt.bin', access='stream', form='unformatted', status='replace')
    write(u) dt
    close(u)
    open(newunit=u, file=trim(dir) // '/out.bin', access='stream', form='unformatted', status='replace')
    write(u) hvy
    close(u)
    print '(A,I0,A)', 'rk4 step of ', nb, ' blocks done'
end program
