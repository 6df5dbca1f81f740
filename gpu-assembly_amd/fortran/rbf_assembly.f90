! rbf_assembly.f90 -- Fortran bind(c) interfaces onto the C ABI of
! librbf_amd (include/rbf_assembly.h), every entry point, plus the few HIP
! runtime calls a Fortran program needs to hold device buffers.
!
! The reference reaches its GPU solve from Fortran through a CUDA-Fortran
! attributes(global) bind(c) interface onto an extern "C" kernel
! (main.f90:6-14,57 -> rbf_operators.cpp:7-11). Here the boundary is a
! plain host-callable C ABI, so any Fortran 2003 compiler binds it with
! ordinary bind(c) interfaces: scalars by value, host arrays by reference,
! device buffers as type(c_ptr) (from hipMalloc of module rbf_hip below, or
! hipfort).
! Arguments that may be NULL in C are type(c_ptr), value: pass c_null_ptr
! or c_loc(array). Status codes are int32, RBFA_OK = 0; C strings come back
! as type(c_ptr) and rbfa_string turns them into Fortran character values.
module rbf_assembly
  use, intrinsic :: iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_float, c_ptr, &
      c_uint64_t, c_char, c_null_char, c_associated, c_f_pointer, c_size_t, c_int
  implicit none
  private

  integer(c_int32_t), parameter, public :: RBFA_OK = 0, RBFA_ERR_ARG = 1, &
      RBFA_ERR_UNSUPPORTED = 2, RBFA_ERR_HIP = 3, RBFA_ERR_NOMEM = 4, RBFA_ERR_NODEVICE = 5, &
      RBFA_ERR_FORMAT = 6, RBFA_ERR_IO = 7
  integer(c_int32_t), parameter, public :: RBFA_OPS_DIFF = 0, RBFA_OPS_INTERP = 1
  integer(c_int32_t), parameter, public :: RBFA_PREC_FP64 = 0, RBFA_PREC_MIXED = 1

  ! rbfa_bbox2, layout of rbf::BBox2<double> (rbf_reorder.h:30-40) and of
  ! the reference's type bbox2d (rbf_ordering.F90:24-26)
  type, bind(c), public :: rbfa_bbox2
    real(c_double) :: xmin, ymin, xmax, ymax
  end type

  type, bind(c), public :: rbfa_plan_desc
    integer(c_int32_t) :: dim, n, degree, phs, op_set, nrhs, precision, max_refine
  end type

  ! library and status text, peaks
  public :: rbfa_version, rbfa_status_string, rbfa_last_error, rbfa_measure_peaks, rbfa_string
  ! host node layer
  public :: rbfa_morton_keys, rbfa_hilbert_keys, rbfa_morton_keys_3d, rbfa_order_by_keys
  public :: rbfa_permute_f64, rbfa_permute_i32, rbfa_renumber_stencils, rbfa_make_row_ptr
  public :: rbfa_knn_stencils, rbfa_gen_jittered_grid, rbfa_gen_graded_disk
  ! device node layer
  public :: rbfa_knn_stencils_device, rbfa_curve_keys_device, rbfa_order_by_keys_device
  public :: rbfa_gather_f64_device
  ! plans and assembly
  public :: rbfa_plan_create, rbfa_plan_destroy, rbfa_plan_nops, rbfa_plan_system_size
  public :: rbfa_assemble, rbfa_assemble_noalloc
  public :: rbfa_plan_num_variants, rbfa_plan_variant_name, rbfa_plan_get_variant
  public :: rbfa_plan_set_variant, rbfa_plan_autotune
  ! uneven batches
  public :: rbfa_batch_create, rbfa_batch_destroy, rbfa_batch_bind, rbfa_batch_class_count
  public :: rbfa_batch_class_variant, rbfa_batch_set_mode, rbfa_batch_get_mode
  public :: rbfa_batch_class_num_variants, rbfa_batch_class_variant_name
  public :: rbfa_batch_set_class_variant
  public :: rbfa_batch_launch, rbfa_batch_autotune
  ! caller-filled systems
  public :: rbfa_gesv_batched, rbfa_solve_dp
  ! operator application
  public :: rbfa_apply, rbfa_apply_plan_create, rbfa_apply_plan_destroy, rbfa_apply_plan_stats
  public :: rbfa_apply_plan_run
  ! files
  public :: rbfa_read_points, rbfa_read_graph, rbfa_write_matrix_market

  interface
    type(c_ptr) function rbfa_version() bind(c)
      import :: c_ptr
    end function

    type(c_ptr) function rbfa_status_string(status) bind(c)
      import :: c_ptr, c_int32_t
      integer(c_int32_t), value :: status
    end function

    type(c_ptr) function rbfa_last_error() bind(c)
      import :: c_ptr
    end function

    ! out(3): FP64 vector FMA TFLOP/s, FP64 MFMA TFLOP/s, HBM copy GB/s
    integer(c_int32_t) function rbfa_measure_peaks(out, stream) bind(c)
      import :: c_int32_t, c_double, c_ptr
      real(c_double), intent(out) :: out(*)
      type(c_ptr), value :: stream
    end function

    integer(c_int32_t) function rbfa_morton_keys(n, x, y, bbox, ndiv, keys) bind(c)
      import :: c_int32_t, c_int64_t, c_double, rbfa_bbox2
      integer(c_int32_t), value :: n, ndiv
      real(c_double), intent(in) :: x(*), y(*)
      type(rbfa_bbox2), intent(in) :: bbox
      integer(c_int64_t), intent(out) :: keys(*)
    end function

    integer(c_int32_t) function rbfa_hilbert_keys(n, x, y, bbox, ndiv, keys) bind(c)
      import :: c_int32_t, c_int64_t, c_double, rbfa_bbox2
      integer(c_int32_t), value :: n, ndiv
      real(c_double), intent(in) :: x(*), y(*)
      type(rbfa_bbox2), intent(in) :: bbox
      integer(c_int64_t), intent(out) :: keys(*)
    end function

    ! bbox6 = {xmin, ymin, zmin, xmax, ymax, zmax}
    integer(c_int32_t) function rbfa_morton_keys_3d(n, x, y, z, bbox6, ndiv, keys) bind(c)
      import :: c_int32_t, c_int64_t, c_double
      integer(c_int32_t), value :: n, ndiv
      real(c_double), intent(in) :: x(*), y(*), z(*), bbox6(6)
      integer(c_int64_t), intent(out) :: keys(*)
    end function

    integer(c_int32_t) function rbfa_order_by_keys(n, keys, perm) bind(c)
      import :: c_int32_t, c_int64_t
      integer(c_int32_t), value :: n
      integer(c_int64_t), intent(in) :: keys(*)
      integer(c_int32_t), intent(out) :: perm(*)
    end function

    integer(c_int32_t) function rbfa_permute_f64(n, perm, a) bind(c)
      import :: c_int32_t, c_double
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: perm(*)
      real(c_double), intent(inout) :: a(*)
    end function

    integer(c_int32_t) function rbfa_permute_i32(n, perm, a) bind(c)
      import :: c_int32_t
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: perm(*)
      integer(c_int32_t), intent(inout) :: a(*)
    end function

    integer(c_int32_t) function rbfa_renumber_stencils(n, k, perm, ja) bind(c)
      import :: c_int32_t
      integer(c_int32_t), value :: n, k
      integer(c_int32_t), intent(in) :: perm(*)
      integer(c_int32_t), intent(inout) :: ja(*)
    end function

    integer(c_int32_t) function rbfa_make_row_ptr(n, k, ia) bind(c)
      import :: c_int32_t
      integer(c_int32_t), value :: n, k
      integer(c_int32_t), intent(out) :: ia(*)
    end function

    ! z may be any array (ignored) when dim == 2
    integer(c_int32_t) function rbfa_knn_stencils(n, dim, x, y, z, k, ja) bind(c)
      import :: c_int32_t, c_double
      integer(c_int32_t), value :: n, dim, k
      real(c_double), intent(in) :: x(*), y(*), z(*)
      integer(c_int32_t), intent(out) :: ja(*)
    end function

    integer(c_int32_t) function rbfa_gen_jittered_grid(dim, m, jitter, seed, x, y, z) bind(c)
      import :: c_int32_t, c_double, c_uint64_t
      integer(c_int32_t), value :: dim, m
      real(c_double), value :: jitter
      integer(c_uint64_t), value :: seed
      real(c_double), intent(out) :: x(*), y(*), z(*)
    end function

    integer(c_int32_t) function rbfa_gen_graded_disk(n_interior, n_boundary, seed, x, y, &
                                                     flag) bind(c)
      import :: c_int32_t, c_double, c_uint64_t
      integer(c_int32_t), value :: n_interior, n_boundary
      integer(c_uint64_t), value :: seed
      real(c_double), intent(out) :: x(*), y(*)
      integer(c_int32_t), intent(out) :: flag(*)
    end function

    ! device node layer: every array argument is a device pointer
    integer(c_int32_t) function rbfa_knn_stencils_device(n, dim, x, y, z, k, ja, stream) bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: n, dim, k
      type(c_ptr), value :: x, y, z, ja, stream
    end function

    integer(c_int32_t) function rbfa_curve_keys_device(n, dim, curve, x, y, z, bbox, ndiv, &
                                                       keys, stream) bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: n, dim, curve, ndiv
      type(c_ptr), value :: x, y, z, bbox, keys, stream
    end function

    integer(c_int32_t) function rbfa_order_by_keys_device(n, keys, key_bits, perm, stream) &
        bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: n, key_bits
      type(c_ptr), value :: keys, perm, stream
    end function

    integer(c_int32_t) function rbfa_gather_f64_device(n, perm, src, dst, stream) bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: n
      type(c_ptr), value :: perm, src, dst, stream
    end function

    integer(c_int32_t) function rbfa_plan_create(desc, plan) bind(c)
      import :: c_int32_t, c_ptr, rbfa_plan_desc
      type(rbfa_plan_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: plan
    end function

    integer(c_int32_t) function rbfa_plan_destroy(plan) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
    end function

    integer(c_int32_t) function rbfa_plan_nops(plan) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
    end function

    integer(c_int32_t) function rbfa_plan_system_size(plan) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
    end function

    ! x, y, z, ja, info, status, xc, yc, zc: device pointers (c_null_ptr
    ! where optional); w: host array of nops device pointers
    integer(c_int32_t) function rbfa_assemble(plan, nstencils, x, y, z, ja, w, xc, yc, zc, &
                                              info, status, stream) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan, x, y, z, ja, xc, yc, zc, info, status, stream
      integer(c_int32_t), value :: nstencils
      type(c_ptr), intent(in) :: w(*)
    end function

    ! rbfa_assemble that allocates nothing (at most 4194304 stencils per call)
    integer(c_int32_t) function rbfa_assemble_noalloc(plan, nstencils, x, y, z, ja, w, xc, yc, &
                                                      zc, info, status, stream) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan, x, y, z, ja, xc, yc, zc, info, status, stream
      integer(c_int32_t), value :: nstencils
      type(c_ptr), intent(in) :: w(*)
    end function

    integer(c_int32_t) function rbfa_plan_num_variants(plan) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
    end function

    type(c_ptr) function rbfa_plan_variant_name(plan, variant) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
      integer(c_int32_t), value :: variant
    end function

    integer(c_int32_t) function rbfa_plan_get_variant(plan) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
    end function

    integer(c_int32_t) function rbfa_plan_set_variant(plan, variant) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
      integer(c_int32_t), value :: variant
    end function

    ! ms_out: c_loc of real(c_float) :: ms(num_variants), or c_null_ptr
    integer(c_int32_t) function rbfa_plan_autotune(plan, nstencils, x, y, z, ja, w, xc, yc, zc, &
                                                   info, status, stream, reps, ms_out) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan, x, y, z, ja, xc, yc, zc, info, status, stream, ms_out
      integer(c_int32_t), value :: nstencils, reps
      type(c_ptr), intent(in) :: w(*)
    end function

    integer(c_int32_t) function rbfa_batch_create(nclasses, classes, batch) bind(c)
      import :: c_int32_t, c_ptr, rbfa_plan_desc
      integer(c_int32_t), value :: nclasses
      type(rbfa_plan_desc), intent(in) :: classes(*)
      type(c_ptr), intent(out) :: batch
    end function

    integer(c_int32_t) function rbfa_batch_destroy(batch) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
    end function

    ! ia_host: host row pointers (nstencils + 1); the rest device pointers,
    ! w a host array of 3 device pointers
    integer(c_int32_t) function rbfa_batch_bind(batch, nstencils, ia_host, x, y, z, ia, ja, w, &
                                                info, status, stream) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch, x, y, z, ia, ja, info, status, stream
      integer(c_int32_t), value :: nstencils
      integer(c_int32_t), intent(in) :: ia_host(*)
      type(c_ptr), intent(in) :: w(*)
    end function

    integer(c_int32_t) function rbfa_batch_class_count(batch, cls) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: cls
    end function

    type(c_ptr) function rbfa_batch_class_variant(batch, cls) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: cls
    end function

    integer(c_int32_t) function rbfa_batch_class_num_variants(batch, cls) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: cls
    end function

    type(c_ptr) function rbfa_batch_class_variant_name(batch, cls, v) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: cls, v
    end function

    integer(c_int32_t) function rbfa_batch_set_class_variant(batch, cls, v) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: cls, v
    end function

    integer(c_int32_t) function rbfa_batch_set_mode(batch, mode) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: mode
    end function

    integer(c_int32_t) function rbfa_batch_get_mode(batch) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int32_t) function rbfa_batch_launch(batch, stream) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch, stream
    end function

    integer(c_int32_t) function rbfa_batch_autotune(batch, stream, reps, ms_out) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch, stream, ms_out
      integer(c_int32_t), value :: reps
    end function

    ! nbatch systems, each A(nt, nt) and B(nt, nrhs) column-major, packed
    ! one after another in device memory; ipiv (nt per system) and info
    ! (one per system) device arrays. A is overwritten by the LU factors,
    ! B by the solution (the reference's execute(A, ipiv, B, info)).
    integer(c_int32_t) function rbfa_gesv_batched(nt, nrhs, nbatch, a, ipiv, b, info, stream) &
        bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: nt, nrhs, nbatch
      type(c_ptr), value :: a, ipiv, b, info, stream
    end function

    ! the reference's solve_dp (rbf_operators.cpp:7-11): one 15 x 15 system,
    ! 9 right-hand sides, device pointers
    integer(c_int32_t) function rbfa_solve_dp(a, ipiv, b, info, stream) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: a, ipiv, b, info, stream
    end function

    ! y_c = W_c u, then y_c(bnd) = u(bnd): the mv_op of src/main.cpp:59-66
    integer(c_int32_t) function rbfa_apply(nrows, k, ia, ja, nops, w, u, y, nbnd, bnd, &
                                           stream) bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: nrows, k, nops, nbnd
      type(c_ptr), value :: ia, ja, u, bnd, stream
      type(c_ptr), intent(in) :: w(*), y(*)
    end function

    ! host arrays: ia (c_null_ptr for fixed width k), ja, bnd
    integer(c_int32_t) function rbfa_apply_plan_create(nrows, ncols, k, ia, ja, nbnd, bnd, &
                                                       plan) bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: nrows, ncols, k, nbnd
      type(c_ptr), value :: ia, ja, bnd
      type(c_ptr), intent(out) :: plan
    end function

    integer(c_int32_t) function rbfa_apply_plan_destroy(plan) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan
    end function

    integer(c_int32_t) function rbfa_apply_plan_stats(plan, out) bind(c)
      import :: c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: plan
      integer(c_int64_t), intent(out) :: out(*)
    end function

    integer(c_int32_t) function rbfa_apply_plan_run(plan, nops, w, u, y, stream) bind(c)
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: plan, u, stream
      integer(c_int32_t), value :: nops
      type(c_ptr), intent(in) :: w(*), y(*)
    end function

    ! path: a NUL-terminated character array (trim(name) // c_null_char);
    ! x, y: c_null_ptr to query the count, else host arrays of cap entries
    integer(c_int32_t) function rbfa_read_points(path, cap, n, x, y) bind(c)
      import :: c_int32_t, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: n
      type(c_ptr), value :: x, y
    end function

    integer(c_int32_t) function rbfa_read_graph(path, k, n, nnz, ia, ja, cap) bind(c)
      import :: c_int32_t, c_int64_t, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: k
      integer(c_int32_t), intent(out) :: n
      integer(c_int64_t), intent(out) :: nnz
      type(c_ptr), value :: ia, ja
      integer(c_int64_t), value :: cap
    end function

    ! a: c_null_ptr writes the pattern only
    integer(c_int32_t) function rbfa_write_matrix_market(path, rows, cols, ia, ja, a, csr_base) &
        bind(c)
      import :: c_int32_t, c_int64_t, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: rows, cols
      integer(c_int32_t), intent(in) :: ia(*), ja(*)
      type(c_ptr), value :: a
      integer(c_int32_t), value :: csr_base
    end function

    function c_strlen(s) bind(c, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! A C string returned by the library as a Fortran character value ('' for
  ! a null pointer).
  function rbfa_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    do i = 1, n
      s(i:i) = chars(i)
    end do
  end function

end module rbf_assembly

! The HIP runtime calls a Fortran caller of librbf_amd needs for device
! buffers (C functions of libamdhip64; status 0 = hipSuccess).
module rbf_hip
  use, intrinsic :: iso_c_binding, only: c_int, c_ptr, c_size_t
  implicit none
  private
  integer(c_int), parameter, public :: hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2
  public :: hipMalloc, hipFree, hipMemcpy, hipDeviceSynchronize

  interface
    integer(c_int) function hipMalloc(ptr, bytes) bind(c, name="hipMalloc")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), intent(out) :: ptr
      integer(c_size_t), value :: bytes
    end function

    integer(c_int) function hipFree(ptr) bind(c, name="hipFree")
      import :: c_int, c_ptr
      type(c_ptr), value :: ptr
    end function

    integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(c, name="hipMemcpy")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int), value :: kind
    end function

    integer(c_int) function hipDeviceSynchronize() bind(c, name="hipDeviceSynchronize")
      import :: c_int
    end function
  end interface
end module rbf_hip
