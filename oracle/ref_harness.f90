! Test harness around the reference's own per-tuple Fortran routines.
!
! TEST INFRASTRUCTURE ONLY.  This program is built by oracle/Makefile from the
! reference sources where they lie (src/common/common_cc.f90 and
! src/triples-ijk/*.f*, plus ZEROT3 from src/triples-abc/ddcc_aux.f) and is used
! to write golden vectors and to time the reference CPU path.  The reference's
! own drivers (cc_triples, cc_triples_acc) need DDI and MPI, which the image
! does not have, so this harness plays the driver role of cc_triples_acc
! (src/triples-ijk/triples.F90:29-320) without any communication: the VVVO
! patches that DDI_GET would return are read from a packed (nutr, nou) array.
!
! Protocol: one stream-unformatted input file, one output file.
!   in : mode, no, nu, i, j, k, lo, hi, eh(no), ep(nu), t1(no*nu), t2(nu*nu*no*no),
!        vm(no*nu*no*no), voe(nu*nu*no*no), packed vvvo(nutr*nu*no)
!   mode 1 : full driver over tuples [lo,hi) (0-based distinct-tuple ids, as
!            DDCC_T_TASK numbers them) plus the degenerate tuples (i,j,j), (i,i,j) of
!            the first k pairs i>j in the driver's order (i outer; k >= no(no-1)/2:
!            all of them); out: ets, etd, v1(nu*no), seconds spent in the tuple loops
!   mode 2 : one distinct tuple (i,j,k) via DDCC_T_IJK_ACC; out: etd, v1, v3(nu^3)
!   mode 3 : one (i,j,j) tuple via DDCC_T_IJJ_ACC; out: etd, v1, v3
!   mode 4 : one (i,i,j) tuple via DDCC_T_IIJ_ACC; out: etd, v1, v3
!   mode 5 : permutation table: the nu^3 cube 1..nu^3 after trant3_1/3/4/5,
!            TRANT3_ACC ids 1..5, tranmd_23, SYMT311_ACC 23/12/13 (13 cubes)
!   mode 6 : VE for occupied index i from the packed patch: DDCC_T_GETVE_acc
!            then trant3_1; out: ve(nu^3)
!   mode 7 : the array a(i,j,k,lo) filled with 1..i*j*k*lo after TRANMD_SMP with
!            code hi; out: a
program ref_harness
use common_cc
implicit none

character(len=512) :: fin, fout
integer :: mode, i, j, k, lo, hi, n, iwrk
integer :: ii, jj, kk
integer(8) :: clk0, clk1, clkrate
double precision, allocatable :: eh(:), ep(:), t1(:), t2(:), vm(:), voe(:), vvvo(:)
double precision, allocatable :: v1(:), v3(:), t3(:), ve_i(:), ve_j(:), ve_k(:), tmp(:), cube(:)
double precision :: ddot
external ddot

call get_command_argument(1, fin)
call get_command_argument(2, fout)

open(10, file=trim(fin), access='stream', form='unformatted', status='old')
read(10) mode, no, nu, i, j, k, lo, hi

! module state, as common_cc_init sets it (src/common/common_cc.f90:342-369),
! for a single process: one SMP rank, one DDI rank
nou = no*nu
no2 = no*no
no3 = no*no2
no4 = no*no3
nu2 = nu*nu
nu3 = nu*nu2
nu4 = nu*nu3
no2u = no2*nu
no3u = no3*nu
nou2 = no*nu2
nou3 = no*nu3
no2u2 = no2*nu2
notr = (no2+no)/2
nutr = (nu2+nu)/2
ets = zero
etd = zero
flops = 0
smp_np = 1
smp_me = 0
ddi_np = 1
ddi_me = 0
ddi_nn = 1
ddi_my = 0

allocate(eh(no), ep(nu), t1(nou), t2(no2u2), vm(no3u), voe(no2u2), vvvo(nutr*nou))
read(10) eh, ep, t1, t2, vm, voe, vvvo
close(10)

allocate(v1(nou), v3(nu3), t3(nu3), ve_i(nu3), ve_j(nu3), ve_k(nu3), tmp(max(nu2, nou)))
v1 = zero
v3 = zero
t3 = zero

open(11, file=trim(fout), access='stream', form='unformatted', status='replace')

select case (mode)
case (1)
   call system_clock(clk0, clkrate)
   do iwrk = lo, hi-1
      call ddcc_t_task(iwrk, no, ii, jj, kk)
      call getve(ii, ve_i)
      call getve(jj, ve_j)
      call getve(kk, ve_k)
      call trant3_1(nu, ve_i)
      call trant3_1(nu, ve_j)
      call trant3_1(nu, ve_k)
      call ddcc_t_ijk_acc(no, nu, ii, jj, kk, v1, t2, vm, v3, voe, eh, ep, ve_i, ve_j, ve_k)
   end do
   if (k .gt. 0) then
      n = 0
      do ii = 1, no
         do jj = 1, ii-1
            n = n + 1
            if (n .gt. k) exit
            call getve(ii, ve_i)
            call tranmd_23(ve_i, nu, nu, nu, 1)
            call getve(jj, ve_j)
            call tranmd_23(ve_j, nu, nu, nu, 1)
            call ddcc_t_ijj_acc(no, nu, ii, jj, v1, t2, vm, v3, t3, voe, t1, eh, ep, tmp, ve_i, ve_j)
            call getve(ii, ve_i)
            call tranmd_23(ve_i, nu, nu, nu, 1)
            call getve(jj, ve_j)
            call tranmd_23(ve_j, nu, nu, nu, 1)
            call ddcc_t_iij_acc(no, nu, ii, jj, v1, t2, vm, v3, t3, voe, t1, eh, ep, tmp, ve_i, ve_j)
         end do
      end do
   end if
   ! final step of cc_triples_acc (src/triples-ijk/triples.F90:310-313)
   call trpose(t1, tmp, no, nu, 1)
   ets = 2.0D+00*ddot(nou, t1, 1, v1, 1)
   call system_clock(clk1)
   write(11) ets, etd, v1, dble(clk1-clk0)/dble(clkrate)
case (2)
   call getve(i, ve_i)
   call getve(j, ve_j)
   call getve(k, ve_k)
   call trant3_1(nu, ve_i)
   call trant3_1(nu, ve_j)
   call trant3_1(nu, ve_k)
   call ddcc_t_ijk_acc(no, nu, i, j, k, v1, t2, vm, v3, voe, eh, ep, ve_i, ve_j, ve_k)
   write(11) etd, v1, v3
case (3)
   call getve(i, ve_i)
   call tranmd_23(ve_i, nu, nu, nu, 1)
   call getve(j, ve_j)
   call tranmd_23(ve_j, nu, nu, nu, 1)
   call ddcc_t_ijj_acc(no, nu, i, j, v1, t2, vm, v3, t3, voe, t1, eh, ep, tmp, ve_i, ve_j)
   write(11) etd, v1, v3
case (4)
   call getve(i, ve_i)
   call tranmd_23(ve_i, nu, nu, nu, 1)
   call getve(j, ve_j)
   call tranmd_23(ve_j, nu, nu, nu, 1)
   call ddcc_t_iij_acc(no, nu, i, j, v1, t2, vm, v3, t3, voe, t1, eh, ep, tmp, ve_i, ve_j)
   write(11) etd, v1, v3
case (5)
   allocate(cube(nu3))
   do n = 1, 13
      call fill(cube)
      select case (n)
      case (1)
         call trant3_1(nu, cube)
      case (2)
         call trant3_3(nu, cube)
      case (3)
         call trant3_4(nu, cube)
      case (4)
         call trant3_5(nu, cube)
      case (5:9)
         call trant3_acc(cube, nu, n-4)
      case (10)
         call tranmd_23(cube, nu, nu, nu, 1)
      case (11)
         call symt311_acc(cube, nu, 23)
      case (12)
         call symt311_acc(cube, nu, 12)
      case (13)
         call symt311_acc(cube, nu, 13)
      end select
      write(11) cube
   end do
case (6)
   call getve(i, ve_i)
   call trant3_1(nu, ve_i)
   write(11) ve_i
case (7)
   n = i*j*k*lo
   allocate(cube(n))
   do iwrk = 1, n
      cube(iwrk) = dble(iwrk)
   end do
   call tranmd_smp(cube, i, j, k, lo, hi)
   write(11) cube
end select
close(11)

contains

  ! The patch DDI_GET(d_vvvo,1,nutr,ilo,ilo+nu-1,tmp) returns for occupied x
  ! (src/triples-ijk/triples.F90:158-162), expanded by the reference's
  ! DDCC_T_GETVE_acc.
  subroutine getve(x, ve)
    integer :: x
    double precision :: ve(nu3)
    integer :: q
    q = 1
    call ddcc_t_getve_acc(q, nu, x, vvvo(1 + nutr*nu*(x-1)), ve)
  end subroutine getve

  subroutine fill(c)
    double precision :: c(nu3)
    integer :: p
    do p = 1, nu3
       c(p) = dble(p)
    end do
  end subroutine fill

end program ref_harness
