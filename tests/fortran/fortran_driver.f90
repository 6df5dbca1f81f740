! fortran_driver.f90 -- a Fortran user of librbf_amd on the GPU, through
! the bind(c) modules of gpu-assembly_amd/fortran/rbf_assembly.f90 only
! (the reference drives its GPU solve from Fortran, main.f90:6-57).
!
!   1. Laplacian / d/dx / d/dy weights of n = 35 stencils (r^5, degree 3)
!      on a 48 x 48 jittered grid: host node layer, then rbfa_assemble
!      on device buffers from hipMalloc.
!   2. The demo's 15 x 15 system with 9 right-hand sides (main.f90:22-57:
!      the 9-point stencil, r^3 + the 6 quadratic monomials, interpolation
!      to the centres (x/2, y/2)), filled here in Fortran, solved by
!      rbfa_solve_dp; then 6 systems derived from it by rbfa_gesv_batched.
!
! Everything the checks need goes to fortran_driver.bin (stream access,
! little-endian): the inputs as filled and the outputs as returned. The
! last line printed is "fortran driver ok".
program fortran_driver
  use, intrinsic :: iso_c_binding
  use rbf_assembly
  use rbf_hip
  implicit none

  integer(c_int32_t), parameter :: m = 48, npts = m * m, ns = 35
  integer(c_int32_t), parameter :: nd = 9, nt = 15, nrhs = 9, nb = 6
  real(c_double), target :: x(npts), y(npts), z(1), w(npts * ns, 3)
  integer(c_int32_t), target :: ja(npts * ns), info(npts)
  real(c_double), target :: a(nt, nt), b(nt, nrhs), a0(nt, nt), b0(nt, nrhs)
  real(c_double), target :: ab(nt, nt, nb), bb(nt, nrhs, nb), ab0(nt, nt, nb), bb0(nt, nrhs, nb)
  integer(c_int32_t), target :: ipiv(nt), info1(1), ipivb(nt, nb), infob(nb)
  type(c_ptr) :: xd, yd, jad, infod, wd(3), plan, ad, bd, ipd, ifd
  type(rbfa_plan_desc) :: desc
  real(c_double) :: px(nd), py(nd)
  integer :: i, j, k, u

  ! ---- 1. assembly through a plan
  call ok(rbfa_gen_jittered_grid(2, m, 0.25d0, 29_c_uint64_t, x, y, z), 'gen_jittered_grid')
  call ok(rbfa_knn_stencils(npts, 2, x, y, z, ns, ja), 'knn_stencils')
  xd = dev_copy(c_loc(x), sizeof(x))
  yd = dev_copy(c_loc(y), sizeof(y))
  jad = dev_copy(c_loc(ja), sizeof(ja))
  call hip(hipMalloc(infod, int(sizeof(info), c_size_t)), 'hipMalloc info')
  do k = 1, 3
    call hip(hipMalloc(wd(k), int(8 * npts * ns, c_size_t)), 'hipMalloc w')
  end do
  desc = rbfa_plan_desc(dim=2, n=ns, degree=3, phs=5, op_set=RBFA_OPS_DIFF, nrhs=0, &
                        precision=RBFA_PREC_FP64, max_refine=3)
  call ok(rbfa_plan_create(desc, plan), 'plan_create')
  if (rbfa_plan_nops(plan) /= 3 .or. rbfa_plan_system_size(plan) /= ns + 10) error stop 'plan'
  print '(a,a)', 'plan variant: ', rbfa_string(rbfa_plan_variant_name(plan, 0))
  call ok(rbfa_assemble(plan, npts, xd, yd, c_null_ptr, jad, wd, c_null_ptr, c_null_ptr, &
                        c_null_ptr, infod, c_null_ptr, c_null_ptr), 'assemble')
  call hip(hipDeviceSynchronize(), 'sync')
  do k = 1, 3
    call host_copy(c_loc(w(1, k)), wd(k), int(8 * npts * ns, c_size_t))
  end do
  call host_copy(c_loc(info), infod, int(sizeof(info), c_size_t))
  call ok(rbfa_plan_destroy(plan), 'plan_destroy')
  print '(a,i0,a,i0)', 'assembled ', npts, ' stencils; info /= 0: ', count(info /= 0)

  ! ---- 2. the demo system, filled here
  px = [0d0, 1d0, 0d0, -1d0, 0d0, 1d0, 1d0, -1d0, -1d0]
  py = [0d0, 0d0, 1d0, 0d0, -1d0, 1d0, -1d0, -1d0, 1d0]
  a = 0d0
  do j = 1, nd
    do i = 1, nd
      a(i, j) = phs3((px(j) - px(i))**2 + (py(j) - py(i))**2)
    end do
    a(nd + 1:nt, j) = mono(px(j), py(j))
    a(j, nd + 1:nt) = mono(px(j), py(j))
  end do
  do j = 1, nrhs
    do i = 1, nd
      b(i, j) = phs3((0.5d0 * px(j) - px(i))**2 + (0.5d0 * py(j) - py(i))**2)
    end do
    b(nd + 1:nt, j) = mono(0.5d0 * px(j), 0.5d0 * py(j))
  end do
  a0 = a
  b0 = b
  ad = dev_copy(c_loc(a), sizeof(a))
  bd = dev_copy(c_loc(b), sizeof(b))
  call hip(hipMalloc(ipd, int(sizeof(ipiv), c_size_t)), 'hipMalloc ipiv')
  call hip(hipMalloc(ifd, int(4 * nb, c_size_t)), 'hipMalloc info')
  call ok(rbfa_solve_dp(ad, ipd, bd, ifd, c_null_ptr), 'solve_dp')
  call hip(hipDeviceSynchronize(), 'sync')
  call host_copy(c_loc(a), ad, int(sizeof(a), c_size_t))
  call host_copy(c_loc(b), bd, int(sizeof(b), c_size_t))
  call host_copy(c_loc(ipiv), ipd, int(sizeof(ipiv), c_size_t))
  call host_copy(c_loc(info1), ifd, 4_c_size_t)
  call hip(hipFree(ad), 'hipFree')
  call hip(hipFree(bd), 'hipFree')
  call hip(hipFree(ipd), 'hipFree')
  print '(a,i0)', 'solve_dp info: ', info1(1)

  ! the batch: system s is the demo's with its diagonal shifted by s/4 and
  ! its right-hand sides scaled by s + 1 (s = 0 .. nb - 1)
  do k = 1, nb
    ab(:, :, k) = a0
    do i = 1, nt
      ab(i, i, k) = ab(i, i, k) + 0.25d0 * (k - 1)
    end do
    bb(:, :, k) = b0 * k
  end do
  ab0 = ab
  bb0 = bb
  ad = dev_copy(c_loc(ab), sizeof(ab))
  bd = dev_copy(c_loc(bb), sizeof(bb))
  call hip(hipMalloc(ipd, int(sizeof(ipivb), c_size_t)), 'hipMalloc ipiv')
  call ok(rbfa_gesv_batched(nt, nrhs, nb, ad, ipd, bd, ifd, c_null_ptr), 'gesv_batched')
  call hip(hipDeviceSynchronize(), 'sync')
  call host_copy(c_loc(ab), ad, int(sizeof(ab), c_size_t))
  call host_copy(c_loc(bb), bd, int(sizeof(bb), c_size_t))
  call host_copy(c_loc(ipivb), ipd, int(sizeof(ipivb), c_size_t))
  call host_copy(c_loc(infob), ifd, int(sizeof(infob), c_size_t))
  print '(a,6(1x,i0))', 'gesv_batched info:', infob

  open(newunit=u, file='fortran_driver.bin', access='stream', form='unformatted', &
       status='replace')
  write(u) x, y, ja, w, info
  write(u) a0, b0, a, b, ipiv, info1
  write(u) ab0, bb0, ab, bb, ipivb, infob
  close(u)
  print '(a)', 'fortran driver ok'

contains

  real(c_double) function phs3(r2)
    real(c_double), intent(in) :: r2
    phs3 = r2 * sqrt(r2)
  end function

  ! 1, x, y, x^2, x y, y^2: degree-major, x powers falling
  function mono(px1, py1) result(v)
    real(c_double), intent(in) :: px1, py1
    real(c_double) :: v(6)
    v = [1d0, px1, py1, px1 * px1, px1 * py1, py1 * py1]
  end function

  subroutine ok(st, what)
    integer(c_int32_t), intent(in) :: st
    character(*), intent(in) :: what
    if (st /= RBFA_OK) then
      print '(a,a,a,a)', what, ': ', rbfa_string(rbfa_status_string(st)), ''
      error stop 1
    end if
  end subroutine

  subroutine hip(st, what)
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: what
    if (st /= 0) then
      print '(a,a,i0)', what, ': hip error ', st
      error stop 2
    end if
  end subroutine

  type(c_ptr) function dev_copy(src, bytes)
    type(c_ptr), intent(in) :: src
    integer(c_size_t), intent(in) :: bytes
    call hip(hipMalloc(dev_copy, bytes), 'hipMalloc')
    call hip(hipMemcpy(dev_copy, src, bytes, hipMemcpyHostToDevice), 'hipMemcpy h2d')
  end function

  subroutine host_copy(dst, src, bytes)
    type(c_ptr), intent(in) :: dst, src
    integer(c_size_t), intent(in) :: bytes
    call hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), 'hipMemcpy d2h')
  end subroutine

end program fortran_driver
