! rbf_assembly.f90 -- Fortran bind(c) interfaces onto the C ABI of
! librbf_amd (include/rbf_assembly.h).
!
! The reference reaches its GPU solve from Fortran through a CUDA-Fortran
! attributes(global) bind(c) interface onto an extern "C" kernel
! (main.f90:6-14,57 -> rbf_operators.cpp:7-11). Here the boundary is a
! plain host-callable C ABI, so any Fortran 2003 compiler binds it with
! ordinary bind(c) interfaces: scalars by value, arrays by reference,
! device buffers as type(c_ptr) obtained from the HIP runtime (hipMalloc
! through hipfort or a C shim). Status codes are int32, RBFA_OK = 0.
module rbf_assembly
  use, intrinsic :: iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_ptr, c_uint64_t
  implicit none
  private

  integer(c_int32_t), parameter, public :: RBFA_OK = 0, RBFA_ERR_ARG = 1, &
      RBFA_ERR_UNSUPPORTED = 2, RBFA_ERR_HIP = 3, RBFA_ERR_NOMEM = 4, RBFA_ERR_NODEVICE = 5
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

  public :: rbfa_morton_keys, rbfa_hilbert_keys, rbfa_morton_keys_3d, rbfa_order_by_keys
  public :: rbfa_renumber_stencils, rbfa_make_row_ptr, rbfa_knn_stencils
  public :: rbfa_gen_jittered_grid, rbfa_plan_create, rbfa_plan_destroy, rbfa_plan_nops
  public :: rbfa_plan_system_size, rbfa_assemble, rbfa_apply

  interface
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

    ! y_c = W_c u, then y_c(bnd) = u(bnd): the mv_op of src/main.cpp:59-66
    integer(c_int32_t) function rbfa_apply(nrows, k, ia, ja, nops, w, u, y, nbnd, bnd, &
                                           stream) bind(c)
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: nrows, k, nops, nbnd
      type(c_ptr), value :: ia, ja, u, bnd, stream
      type(c_ptr), intent(in) :: w(*), y(*)
    end function
  end interface

end module rbf_assembly
