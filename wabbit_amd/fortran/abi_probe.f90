! Links a Fortran program against libwabbit_hip.so through the bind(C)
! interface and calls the entry points that need no GPU.
program abi_probe
    use wabbit_hip
    implicit none
    integer(c_int) :: v
    type(wh_pool_desc) :: d
    v = wh_version()
    print '(A,I0)', 'wh_version=', v
    print '(A,I0)', 'sizeof(wh_pool_desc)=', c_sizeof(d)
end program
