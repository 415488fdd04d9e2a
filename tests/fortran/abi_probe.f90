! Links a Fortran program against libwabbit_hip.so through the bind(C)
! interface and calls the entry points that need no GPU.
program abi_probe
    use wabbit_hip
    implicit none
    integer(c_int) :: v, n
    type(wh_pool_desc) :: d
    type(wh_wavelet) :: w
    type(wh_adapt_params) :: ap
    type(wh_mask_params) :: mp
    type(wh_time_params) :: tp
    type(c_ptr) :: forest
    integer(c_int) :: bs(3), per(3)
    real(c_double) :: dom(3)
    integer(c_int64_t) :: tc(4)
    integer(c_int32_t) :: lv(4), hv(4)
    v = wh_version()
    print '(A,I0)', 'wh_version=', v
    print '(A,I0)', 'sizeof(wh_pool_desc)=', c_sizeof(d)
    print '(A,I0)', 'sizeof(wh_wavelet)=', c_sizeof(w)
    print '(A,I0)', 'sizeof(wh_adapt_params)=', c_sizeof(ap)
    print '(A,I0)', 'sizeof(wh_mask_params)=', c_sizeof(mp)
    print '(A,I0)', 'sizeof(wh_time_params)=', c_sizeof(tp)
    call wh_check(wh_wavelet_setup("CDF44" // c_null_char, 2, w), "setup_wavelet")
    print '(A,I0,A,I0,A,I0)', 'CDF44 g=', w%g, ' g_rhs=', w%g_rhs, ' lifted=', w%lifted
    bs = [16, 16, 1]
    per = [1, 1, 1]
    dom = [1.0d0, 1.0d0, 1.0d0]
    call wh_check(wh_forest_create(forest, 2, 6, bs, 6, dom, per, 64), "forest_create")
    ! level 1: digits 0..3 at bit offset (jmax-1)*dim = 10
    tc = [0_c_int64_t, 1024_c_int64_t, 2048_c_int64_t, 3072_c_int64_t]
    lv = 1
    hv = [0, 1, 2, 3]
    call wh_check(wh_forest_set_blocks(forest, 4, tc, lv, hv), "forest_set_blocks")
    call wh_check(wh_forest_size(forest, n), "forest_size")
    print '(A,I0)', 'forest_n=', n
    call wh_check(wh_forest_destroy(forest), "forest_destroy")
end program
