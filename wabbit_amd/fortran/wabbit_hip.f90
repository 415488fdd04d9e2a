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

    integer(c_int), parameter :: WH_WAVELET_TAPS = 33
    integer(c_int), parameter :: WH_IND_EVERYWHERE = 0, WH_IND_THRESHOLD = 1, WH_IND_EXTERNAL = 2
    integer(c_int), parameter :: WH_NORM_LINFTY = 0, WH_NORM_L1 = 1, WH_NORM_L2 = 2, WH_NORM_H1 = 3

    type, bind(C) :: wh_wavelet
        integer(c_int) :: X, Y, lifted, specialised
        integer(c_int) :: hd_lo, hd_hi, gd_lo, gd_hi, hr_lo, hr_hi, gr_lo, gr_hi
        real(c_double) :: HD(WH_WAVELET_TAPS), GD(WH_WAVELET_TAPS), HR(WH_WAVELET_TAPS), GR(WH_WAVELET_TAPS)
        integer(c_int) :: g, g_rhs, order
        integer(c_int) :: Nscl, Nscr, Nwcl, Nwcr, Nreconl, Nreconr, Ndep2l, Ndep2r
    end type

    type, bind(C) :: wh_mask_params
        integer(c_int) :: geometry
        real(c_double) :: x_cntr(3), R_cyl
        real(c_double) :: smoothing_width, smoothing_safety
        integer(c_int) :: use_sponge, sponge_type
        real(c_double) :: L_sponge, p_sponge
        real(c_double) :: domain(3)
    end type

    type, bind(C) :: wh_adapt_params
        integer(c_int) :: jmin, jmax, indicator, use_ce, force_maxlevel_dealiasing
        real(c_double) :: eps
        real(c_double) :: norm(8)
        integer(c_int32_t) :: groups(8)
        integer(c_int) :: eps_normalized, threshold_mask, security_zone
        real(c_double) :: domain(3)
        type(wh_mask_params) :: mask
        integer(c_int) :: decompose_branch
        integer(c_int) :: eps_norm
    end type

    type, bind(C) :: wh_time_params
        real(c_double) :: time_max, dt_fixed, dt_max
        real(c_double) :: write_time, write_time_first, tsave_stats, write_backup_time
        integer(c_int) :: write_fixed_time
        integer(c_int) :: write_backup_n
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
        ! ---- adaptive path: forest, wavelets, mask, adapt_tree ----
        function wh_forest_create(forest, dim, jmax, bs, g, domain, periodic, nb_max) &
                bind(C, name="wh_forest_create") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), intent(out) :: forest
            integer(c_int), value :: dim, jmax, g, nb_max
            integer(c_int), intent(in) :: bs(3), periodic(3)
            real(c_double), intent(in) :: domain(3)
            integer(c_int) :: rc
        end function
        function wh_forest_destroy(forest) bind(C, name="wh_forest_destroy") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: forest
            integer(c_int) :: rc
        end function
        function wh_forest_set_blocks(forest, n, treecodes, levels, hvy_ids) &
                bind(C, name="wh_forest_set_blocks") result(rc)
            import :: c_ptr, c_int, c_int32_t, c_int64_t
            type(c_ptr), value :: forest
            integer(c_int), value :: n
            integer(c_int64_t), intent(in) :: treecodes(*)
            integer(c_int32_t), intent(in) :: levels(*), hvy_ids(*)
            integer(c_int) :: rc
        end function
        function wh_forest_size(forest, n) bind(C, name="wh_forest_size") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: forest
            integer(c_int), intent(out) :: n
            integer(c_int) :: rc
        end function
        function wh_forest_get_blocks(forest, cap, treecodes, levels, hvy_ids, x0dx) &
                bind(C, name="wh_forest_get_blocks") result(rc)
            import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: forest
            integer(c_int), value :: cap
            integer(c_int64_t), intent(out) :: treecodes(*)
            integer(c_int32_t), intent(out) :: levels(*), hvy_ids(*)
            real(c_double), intent(out) :: x0dx(*)
            integer(c_int) :: rc
        end function
        function wh_wavelet_setup(name, fd_max_size, w) bind(C, name="wh_wavelet_setup") result(rc)
            import :: c_int, c_char, wh_wavelet
            character(kind=c_char), intent(in) :: name(*)
            integer(c_int), value :: fd_max_size
            type(wh_wavelet), intent(out) :: w
            integer(c_int) :: rc
        end function
        function wh_pool_set_wavelet(pool, w) bind(C, name="wh_pool_set_wavelet") result(rc)
            import :: c_ptr, c_int, wh_wavelet
            type(c_ptr), value :: pool
            type(wh_wavelet), intent(in) :: w
            integer(c_int) :: rc
        end function
        function wh_create_mask(pool, mp, hvy_ids, n, x0dx) bind(C, name="wh_create_mask") result(rc)
            import :: c_ptr, c_int, wh_mask_params
            type(c_ptr), value :: pool
            type(wh_mask_params), intent(in) :: mp
            type(c_ptr), value :: hvy_ids, x0dx
            integer(c_int), value :: n
            integer(c_int) :: rc
        end function
        function wh_sync_tree(pool, forest, array) bind(C, name="wh_sync_tree") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: pool, forest
            integer(c_int), value :: array
            integer(c_int) :: rc
        end function
        function wh_adapt_tree(pool, forest, p) bind(C, name="wh_adapt_tree") result(rc)
            import :: c_ptr, c_int, wh_adapt_params
            type(c_ptr), value :: pool, forest
            type(wh_adapt_params), intent(in) :: p
            integer(c_int) :: rc
        end function
        function wh_adapt_tree_flags(pool, forest, p, leaf_flags, n) bind(C, name="wh_adapt_tree_flags") result(rc)
            import :: c_ptr, c_int, c_int32_t, wh_adapt_params
            type(c_ptr), value :: pool, forest
            type(wh_adapt_params), intent(in) :: p
            integer(c_int32_t), intent(in) :: leaf_flags(*)
            integer(c_int), value :: n
            integer(c_int) :: rc
        end function
        function wh_fwt_threshold(pool, src_array, dst_array, hvy_ids, n, thresh, groups, flags_out, detail_out) &
                bind(C, name="wh_fwt_threshold") result(rc)
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: pool
            integer(c_int), value :: src_array, dst_array, n
            integer(c_int32_t), intent(in) :: hvy_ids(*), groups(*)
            real(c_double), intent(in) :: thresh(*)
            integer(c_int32_t), intent(out) :: flags_out(*)
            real(c_double), intent(out) :: detail_out(*)
            integer(c_int) :: rc
        end function
        function wh_detail_blocks_norm(pool, array, hvy_ids, levels, n, groups, eps_norm, detail) &
                bind(C, name="wh_detail_blocks_norm") result(rc)
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: pool
            integer(c_int), value :: array, n, eps_norm
            integer(c_int32_t), intent(in) :: hvy_ids(*), levels(*), groups(*)
            real(c_double), intent(out) :: detail(*)
            integer(c_int) :: rc
        end function
        function wh_component_norm(pool, array, hvy_ids, levels, n, groups, eps_norm, domain, norm) &
                bind(C, name="wh_component_norm") result(rc)
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: pool
            integer(c_int), value :: array, n, eps_norm
            integer(c_int32_t), intent(in) :: hvy_ids(*), levels(*), groups(*)
            real(c_double), intent(in) :: domain(3)
            real(c_double), intent(out) :: norm(*)
            integer(c_int) :: rc
        end function
        function wh_forest_refine(forest, n, hvy_ids, triples, cap, ntriples) &
                bind(C, name="wh_forest_refine") result(rc)
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: forest
            integer(c_int), value :: n, cap
            integer(c_int32_t), intent(in) :: hvy_ids(*)
            integer(c_int32_t), intent(out) :: triples(*)
            integer(c_int), intent(out) :: ntriples
            integer(c_int) :: rc
        end function
        function wh_forest_set_slot_base(forest, first) bind(C, name="wh_forest_set_slot_base") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: forest
            integer(c_int), value :: first
            integer(c_int) :: rc
        end function
        function wh_set_block_groups(pool, group, n, ngroups) bind(C, name="wh_set_block_groups") result(rc)
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: pool
            integer(c_int32_t), intent(in) :: group(*)
            integer(c_int), value :: n, ngroups
            integer(c_int) :: rc
        end function
        function wh_dt_reduce_groups(pool, p, dt_out) bind(C, name="wh_dt_reduce_groups") result(rc)
            import :: c_ptr, c_int, c_double, wh_acm_params
            type(c_ptr), value :: pool
            type(wh_acm_params), intent(in) :: p
            real(c_double), intent(out) :: dt_out(*)
            integer(c_int) :: rc
        end function
        function wh_rk_step_groups(pool, p, butcher, nrows, dt_groups, path) bind(C, name="wh_rk_step_groups") result(rc)
            import :: c_ptr, c_int, c_double, wh_acm_params
            type(c_ptr), value :: pool
            type(wh_acm_params), intent(in) :: p
            real(c_double), intent(in) :: butcher(*), dt_groups(*)
            integer(c_int), value :: nrows, path
            integer(c_int) :: rc
        end function
        function wh_time_set(pool, time, iteration) bind(C, name="wh_time_set") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: pool
            real(c_double), value :: time
            integer(c_int), value :: iteration
            integer(c_int) :: rc
        end function
        function wh_run_steps(pool, p, butcher, nrows, tp, nsteps, use_graph) bind(C, name="wh_run_steps") result(rc)
            import :: c_ptr, c_int, c_double, wh_acm_params, wh_time_params
            type(c_ptr), value :: pool
            type(wh_acm_params), intent(in) :: p
            real(c_double), intent(in) :: butcher(*)
            integer(c_int), value :: nrows, nsteps, use_graph
            type(wh_time_params), intent(in) :: tp
            integer(c_int) :: rc
        end function
        function wh_time_fetch(pool) bind(C, name="wh_time_fetch") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: pool
            integer(c_int) :: rc
        end function
        function wh_time_fetched(pool, time, iteration, dt_last) bind(C, name="wh_time_fetched") result(rc)
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: pool
            real(c_double), intent(out) :: time, dt_last
            integer(c_int), intent(out) :: iteration
            integer(c_int) :: rc
        end function
        function wh_refine_blocks(pool, array, triples, ntriples, order) bind(C, name="wh_refine_blocks") result(rc)
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: pool
            integer(c_int), value :: array, ntriples, order
            integer(c_int32_t), intent(in) :: triples(*)
            integer(c_int) :: rc
        end function
        ! several forests sharing one pool (forests(1:n): wh_forest handles)
        function wh_sync_trees(pool, forests, n, array) bind(C, name="wh_sync_trees") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: pool
            type(c_ptr), intent(in) :: forests(*)
            integer(c_int), value :: n, array
            integer(c_int) :: rc
        end function
        function wh_adapt_trees(pool, forests, n, params, status) bind(C, name="wh_adapt_trees") result(rc)
            import :: c_ptr, c_int, wh_adapt_params
            type(c_ptr), value :: pool
            type(c_ptr), intent(in) :: forests(*)
            integer(c_int), value :: n
            type(wh_adapt_params), intent(in) :: params(*)
            integer(c_int), intent(out) :: status(*)
            integer(c_int) :: rc
        end function
        function wh_refine_trees(pool, forests, n, jmax, order, array, status) bind(C, name="wh_refine_trees") result(rc)
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: pool
            type(c_ptr), intent(in) :: forests(*)
            integer(c_int), value :: n, order, array
            integer(c_int32_t), intent(in) :: jmax(*)
            integer(c_int), intent(out) :: status(*)
            integer(c_int) :: rc
        end function
        function wh_create_masks(pool, forests, n, params, status) bind(C, name="wh_create_masks") result(rc)
            import :: c_ptr, c_int, wh_mask_params
            type(c_ptr), value :: pool
            type(c_ptr), intent(in) :: forests(*)
            integer(c_int), value :: n
            type(wh_mask_params), intent(in) :: params(*)
            integer(c_int), intent(out) :: status(*)
            integer(c_int) :: rc
        end function
        function wh_set_tables_from_forests(pool, forests, n, g_rhs, order, status, nblocks) &
                bind(C, name="wh_set_tables_from_forests") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: pool
            type(c_ptr), intent(in) :: forests(*)
            integer(c_int), value :: n, g_rhs, order
            integer(c_int), intent(out) :: status(*), nblocks(*)
            integer(c_int) :: rc
        end function
        function wh_set_group_params(pool, params, ngroups) bind(C, name="wh_set_group_params") result(rc)
            import :: c_ptr, c_int, wh_acm_params
            type(c_ptr), value :: pool
            type(wh_acm_params), intent(in) :: params(*)
            integer(c_int), value :: ngroups
            integer(c_int) :: rc
        end function
        ! RungeKuttaGeneric over several pools of one dimension at once
        function wh_batch_create(batch, pools, npools) bind(C, name="wh_batch_create") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), intent(out) :: batch
            type(c_ptr), intent(in) :: pools(*)
            integer(c_int), value :: npools
            integer(c_int) :: rc
        end function
        function wh_batch_destroy(batch) bind(C, name="wh_batch_destroy") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: batch
            integer(c_int) :: rc
        end function
        function wh_batch_rk_step(batch, p, butcher, nrows, dt_groups, use_graph) &
                bind(C, name="wh_batch_rk_step") result(rc)
            import :: c_ptr, c_int, c_double, wh_acm_params
            type(c_ptr), value :: batch
            type(wh_acm_params), intent(in) :: p
            real(c_double), intent(in) :: butcher(*), dt_groups(*)
            integer(c_int), value :: nrows, use_graph
            integer(c_int) :: rc
        end function
        function wh_batch_synchronize(batch) bind(C, name="wh_batch_synchronize") result(rc)
            import :: c_ptr, c_int
            type(c_ptr), value :: batch
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
