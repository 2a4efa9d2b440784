! A Fortran caller of libcc_mi355x, built against the library the way a reference caller links
! against libcc (8-byte default INTEGER, src/triples-ijk/Makefile:10).  Two routes:
!   1. the GAMESS-facing legacy entry, as the reference's callers use it:
!      common_cc_init (src/common/common_cc.f90:342), then
!      cc_triples_acc(eh,ep,v1,t1,t2,v3,t3,vm,voe,vvvo_hnd,gms_ets,gms_etd)
!      (src/triples-ijk/triples.F90:1-12), with VVVO registered behind the handle;
!   2. the context API through an iso_c_binding interface block (ccx_create, ccx_load_host,
!      ccx_energy, ccx_destroy).
! Input (stream file): no, nu (int64), then eh, ep, t1(no,nu), t2, vm, voe, vvvo(nu,nu,nu,no).
! Output (stream file): status of the legacy call, gms_ets, gms_etd, status of the context
! calls, ets, etd, max |t1 after the call - transpose(t1 before)|, max |v1 legacy - v1 context|.
module ccx_iface
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: CCX_FORM_FP64 = 0
  interface
    integer(c_int) function ccx_create(no, nu, ctx) bind(C, name="ccx_create")
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: no, nu
      type(c_ptr) :: ctx
    end function
    integer(c_int) function ccx_destroy(ctx) bind(C, name="ccx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ccx_load_host(ctx, eh, ep, t1, t2, vm, voe, vvvo, layout) bind(C, name="ccx_load_host")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: eh(*), ep(*), t1(*), t2(*), vm(*), voe(*), vvvo(*)
      integer(c_int), value :: layout
    end function
    integer(c_int) function ccx_energy(ctx, form, ets, etd, v1, v1_on_device) bind(C, name="ccx_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: form
      real(c_double) :: ets, etd
      real(c_double) :: v1(*)
      integer(c_int), value :: v1_on_device
    end function
    integer(c_int) function ccx_last_status() bind(C, name="ccx_last_status")
      import :: c_int
    end function
  end interface
end module ccx_iface

program ccx_caller
  use ccx_iface
  implicit none
  external common_cc_init, ccx_register_vvvo, cc_triples_acc
  character(len=1024) :: fin, fout
  integer :: no, nu, hnd, layout, i, a
  integer(c_int) :: st_legacy, st_ctx
  double precision, allocatable :: eh(:), ep(:), t1(:,:), t1w(:), t2(:), vm(:), voe(:), vvvo(:)
  double precision, allocatable :: v1(:), v1c(:), v3(:), t3(:)
  double precision :: gms_ets, gms_etd, ets, etd, tdiff
  type(c_ptr) :: ctx

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) no, nu
  allocate(eh(no), ep(nu), t1(no, nu), t1w(no*nu), t2(nu*nu*no*no), vm(no*nu*no*no), voe(nu*nu*no*no))
  allocate(vvvo(nu*nu*nu*no), v1(nu*no), v1c(nu*no), v3(nu*nu*nu), t3(nu*nu*nu))
  read(10) eh, ep, t1, t2, vm, voe, vvvo
  close(10)

  ! 1. legacy ABI: t1 is passed as the working copy the driver transposes in place
  t1w = reshape(t1, (/ no*nu /))
  v1 = 0.0d0
  call common_cc_init(no, nu)
  hnd = 17
  layout = 0
  call ccx_register_vvvo(hnd, vvvo, layout)
  call cc_triples_acc(eh, ep, v1, t1w, t2, v3, t3, vm, voe, hnd, gms_ets, gms_etd)
  st_legacy = ccx_last_status()
  tdiff = 0.0d0
  do i = 1, no
    do a = 1, nu
      tdiff = max(tdiff, abs(t1w(a + nu*(i-1)) - t1(i, a)))   ! t1w is now (nu,no)
    end do
  end do

  ! 2. context API through iso_c_binding
  st_ctx = ccx_create(int(no, c_int64_t), int(nu, c_int64_t), ctx)
  if (st_ctx == 0) st_ctx = ccx_load_host(ctx, eh, ep, t1, t2, vm, voe, vvvo, 0_c_int)
  ets = 0.0d0
  etd = 0.0d0
  v1c = 0.0d0
  if (st_ctx == 0) st_ctx = ccx_energy(ctx, CCX_FORM_FP64, ets, etd, v1c, 0_c_int)
  if (c_associated(ctx)) i = ccx_destroy(ctx)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) dble(st_legacy), gms_ets, gms_etd, dble(st_ctx), ets, etd, tdiff, maxval(abs(v1 - v1c))
  close(11)
end program ccx_caller
