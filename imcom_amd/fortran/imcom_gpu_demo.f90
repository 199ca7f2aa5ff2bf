! Fortran driver of the GPU solve path through the ISO_C_BINDING module:
! the reference's get_A / get_QL / get_B / get_P / get_T / build_image
! sequence (imcom.f90:151-169) on one tile read from a stream file, with the
! outputs written back.  It shows the Fortran side of the C ABI working as
! the reference driver would use it.
!
!   imcom_gpu_demo <input.bin> <output.bin>
!
! input  (stream, native endianness): int64 n, m; int32 psfconst, nexppsf,
!        n1, n2, npoly, npad, maxnits, UnotS; real64 C_a, kappa_min,
!        kappa_max, US_max, US_tol, psfxscale, psfyscale; then x_i(n), y_i(n),
!        exp_i(n) (int32), X_a(m), Y_a(m), Ndiag(n), I_i(n),
!        Alookup(n1*n2*nexppsf*(nexppsf+1)/2), Blookup(n1*n2*nexppsf)
! output: K(m), U(m), S(m), H(m), condition
PROGRAM imcom_gpu_demo
  USE, INTRINSIC :: iso_c_binding
  USE imcom_gpu_iface
  implicit none
  type(c_ptr) :: ctx
  integer(c_int64_t) :: n, m
  integer(c_int32_t) :: psfconst, nexppsf, n1, n2, npoly, npad, maxnits, unots
  real(c_double) :: C_a, kmin, kmax, usmax, ustol, sx, sy, cond
  real(c_double), allocatable :: x(:), y(:), Xa(:), Ya(:), Nd(:), Ii(:), Atab(:), Btab(:)
  real(c_double), allocatable :: K(:), US(:), SU(:), H(:), L(:)
  integer(c_int32_t), allocatable :: ex(:)
  integer(c_int) :: st
  integer :: nA, nB
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) n, m
  read(10) psfconst, nexppsf, n1, n2, npoly, npad, maxnits, unots
  read(10) C_a, kmin, kmax, usmax, ustol, sx, sy
  nA = merge(1, nexppsf * (nexppsf + 1) / 2, psfconst.eq.1)
  nB = merge(1, nexppsf, psfconst.eq.1)
  allocate(x(n), y(n), ex(n), Xa(m), Ya(m), Nd(n), Ii(n))
  allocate(Atab(int(n1, 8) * n2 * nA), Btab(int(n1, 8) * n2 * nB))
  read(10) x, y, ex, Xa, Ya, Nd, Ii, Atab, Btab
  close(10)

  st = imcom_gpu_init(ctx, 0_c_int)
  if (st.ne.0) then
    write(*, FMT='(A)') "IMCOM ERROR: no usable GPU"
    stop
  end if
  call imcom_gpu_check(ctx, imcom_gpu_set_tile(ctx, n, m, x, y, ex, Xa, Ya, Nd, C_a), "set_tile")
  call imcom_gpu_check(ctx, imcom_gpu_set_tables(ctx, int(psfconst, c_int), int(nexppsf, c_int), &
                       int(n1, c_int), int(n2, c_int), Atab, Btab), "set_tables")
  allocate(L(n))
  allocate(K(m), US(m), SU(m), H(m))
  ! imcom_get_A -> imcom_lookup_A (imcom_matrix.f90:206); A stays on the device
  call imcom_gpu_check(ctx, imcom_gpu_lookup_A(ctx, int(npoly, c_int), int(npad, c_int), sx, sy, &
                       c_null_ptr), "lookup_A")
  ! imcom_get_QL -> imcom_eigen_A (imcom_matrix.f90:586); Q stays on the device
  call imcom_gpu_check(ctx, imcom_gpu_eigen_A(ctx, c_null_ptr, imcom_gpu_addr(L), cond), "eigen_A")
  write(*, FMT='(A, E24.16)') "IMCOM: A matrix condition number = ", cond
  ! imcom_get_B -> imcom_lookup_B (imcom_matrix.f90:288)
  call imcom_gpu_check(ctx, imcom_gpu_lookup_B(ctx, int(npoly, c_int), int(npad, c_int), sx, sy, &
                       c_null_ptr), "lookup_B")
  ! imcom_get_P -> imcom_build_P (imcom_matrix.f90:682)
  call imcom_gpu_check(ctx, imcom_gpu_build_P(ctx, IMCOM_PREC_FP64, c_null_ptr), "build_P")
  ! imcom_get_T -> imcom_bisect_kappa + imcom_build_T (imcom_bisect.f90:40-74)
  call imcom_gpu_check(ctx, imcom_gpu_bisect_kappa(ctx, int(unots, c_int), usmax, kmin, kmax, ustol, &
                       int(maxnits, c_int), c_null_ptr, K, US, c_null_ptr), "bisect_kappa")
  call imcom_gpu_check(ctx, imcom_gpu_build_T(ctx, IMCOM_PREC_FP64, c_null_ptr), "build_T")
  if (unots.ne.0) then
    call imcom_gpu_check(ctx, imcom_gpu_build_S(ctx, SU), "build_S")
  else
    call imcom_gpu_check(ctx, imcom_gpu_build_U(ctx, SU), "build_U")
  end if
  ! imcom_build_image (imcom_matrix.f90:754)
  call imcom_gpu_check(ctx, imcom_gpu_build_image(ctx, Ii, H), "build_image")
  call imcom_gpu_finalize(ctx)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  if (unots.ne.0) then
    write(11) K, US, SU, H, cond
  else
    write(11) K, SU, US, H, cond
  end if
  close(11)
  write(*, FMT='(A)') "IMCOM: GPU solve path complete"
END PROGRAM imcom_gpu_demo
