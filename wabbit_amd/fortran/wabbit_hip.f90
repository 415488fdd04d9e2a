! wabbit_hip.f90 -- ISO_C_BINDING interface of libwabbit_hip.so for a Fortran
! host (the C declarations are in include/wabbit_hip.h).  Heavy-data arrays
! keep the host's layout hvy_block(nx, ny, nz, neq, Nb); hvy ids are 0-based
! on this side of the boundary (hvy_id - 1 in WABBIT's numbering).
module wabbit_hip
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: WH_OK = 0, WH_BLOCK = 0, WH_TMP = 1, WH_MASK = 2, WH_WORK = 16
    integer(c_int), parameter :: WH_FD2 = 2, WH_FD4 = 4, WH_FD4_TW = 40, WH_FD6 = 6

    type, bind(C) :: wh_pool_desc
        integer(c_int) :: dim
        integer(c_int) :: bs(3)
        integer(c_int) :: g, neq, nb_max, n_work_slots, n_tmp_eqn, n_mask, mask_per_block
    end type

    type, bind(C) :: wh_acm_params
        integer(c_int) :: dim, scheme, skew_symmetry, penalization, use_sponge
        real(c_double) :: c_0, nu, gamma_p, C_eta, C_sponge
        real(c_double) :: u_mean_set(3)
        real(c_double) :: CFL, CFL_nu, CFL_eta
    end type

    type, bind(C) :: wh_patch
        integer(c_int32_t) :: recv, send
        integer(c_int32_t) :: recv_lo(3), send_lo(3), n(3)
        integer(c_int32_t) :: op, stage, aux
    end type

    interface
        function wh_version() bind(C, name="wh_version") result(v)
            import :: c_int
            integer(c_int) :: v
        end function
        function wh_last_error() bind(C, name="wh_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        function wh_pool_create(pool, desc) bind(C, name="wh_pool_create") result(rc)
            import :: c_ptr, c_int, wh_pool_desc
            type(c_ptr), intent(out) :: pool
            type(wh_pool_desc), intent(in) :: desc
            integer(c_int) :: rc
        end function
        function wh_pool_destroy(pool) bind(C, name="wh_pool_destroy") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: pool
            integer(c_int) :: rc
        end function
        function wh_pool_upload(pool, array, first, count, host) bind(C, name="wh_pool_upload") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: pool
            integer(c_int), value :: array, first, count
            real(c_double), intent(in) :: host(*)
            integer(c_int) :: rc
        end function
        function wh_pool_download(pool, array, first, count, host) bind(C, name="wh_pool_download") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: pool
            integer(c_int), value :: array, first, count
            real(c_double), intent(out) :: host(*)
            integer(c_int) :: rc
        end function
        function wh_set_active(pool, hvy_active, n, x0dx) bind(C, name="wh_set_active") result(rc)
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: pool
            integer(c_int32_t), intent(in) :: hvy_active(*)
            integer(c_int), value :: n
            real(c_double), intent(in) :: x0dx(*)
            integer(c_int) :: rc
        end function
        function wh_set_patches(pool, sync_case, patches, npatch) bind(C, name="wh_set_patches") result(rc)
            import :: c_ptr, c_int, wh_patch
            type(c_ptr), value :: pool
            integer(c_int), value :: sync_case, npatch
            type(wh_patch), intent(in) :: patches(*)
            integer(c_int) :: rc
        end function
        function wh_sync_ghosts(pool, array, sync_case) bind(C, name="wh_sync_ghosts") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: pool
            integer(c_int), value :: array, sync_case
            integer(c_int) :: rc
        end function
        function wh_rhs_acm(pool, p, time, in_array, out_slot) bind(C, name="wh_rhs_acm") result(rc)
            import :: c_ptr, c_int, c_double, wh_acm_params
            type(c_ptr), value :: pool
            type(wh_acm_params), intent(in) :: p
            real(c_double), value :: time
            integer(c_int), value :: in_array, out_slot
            integer(c_int) :: rc
        end function
        function wh_dt_reduce(pool, p, dt) bind(C, name="wh_dt_reduce") result(rc)
            import :: c_ptr, c_int, c_double, wh_acm_params
            type(c_ptr), value :: pool
            type(wh_acm_params), intent(in) :: p
            real(c_double), intent(out) :: dt
            integer(c_int) :: rc
        end function
        function wh_rk_step(pool, p, butcher, nrows, time, dt, path) bind(C, name="wh_rk_step") result(rc)
            import :: c_ptr, c_int, c_double, wh_acm_params
            type(c_ptr), value :: pool
            type(wh_acm_params), intent(in) :: p
            real(c_double), intent(in) :: butcher(*)
            integer(c_int), value :: nrows, path
            real(c_double), value :: time, dt
            integer(c_int) :: rc
        end function
    end interface

contains

    ! the reference aborts on errors (module_globals.f90:131-162); do the same
    subroutine wh_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc == WH_OK) return
        call c_f_pointer(wh_last_error(), msg, [1024])
        n = 0
        do while (n < 1024)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        end do
        write(*, '(A,": wabbit_hip error ",I0,": ",1024A)') where, rc, msg(1:n)
        error stop 1
    end subroutine

end module wabbit_hip
