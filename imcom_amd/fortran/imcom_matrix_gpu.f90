! MODULE imcom_matrix for the MI355X solve path: a drop-in for the reference's
! imcom_matrix.f90.  The driver imcom.f90 and the reference's imcom_data,
! imcom_io and imcom_proc modules compile unchanged against it; every public
! procedure the driver and imcom_bisect call keeps its name and arguments:
!
!   imcom_get_A(npoly, npad, saveA, forcebuild)   imcom_matrix.f90:41-74
!   imcom_get_B(npoly, npad, saveB, forcebuild)   imcom_matrix.f90:77-110
!   imcom_get_P(saveP, forcebuild)                imcom_matrix.f90:113-152
!   imcom_get_QL(saveQL, forcebuild)              imcom_matrix.f90:155-202
!   imcom_lookup_A(npoly, npad, saveA)            imcom_matrix.f90:206-284
!   imcom_lookup_B(npoly, npad, saveB)            imcom_matrix.f90:288-357
!   imcom_build_Alookup(npad)                     imcom_matrix.f90:361-430
!   imcom_build_Blookup(npad)                     imcom_matrix.f90:434-493
!   imcom_calc_C                                  imcom_matrix.f90:546-552
!   imcom_build_Ndiag                             imcom_matrix.f90:556-582
!   imcom_eigen_A(saveQL)                         imcom_matrix.f90:586-678
!   imcom_build_P(saveP)                          imcom_matrix.f90:682-712
!   imcom_build_T                                 imcom_matrix.f90:716-750
!   imcom_build_image                             imcom_matrix.f90:754-790
!   imcom_build_U / imcom_build_S                 imcom_matrix.f90:795-877
!
! The arithmetic runs on the GPU through libimcom_hip.so (imcom_gpu_iface).
! One C context per process is opened at the first call; the tile's input
! vectors (x_i, y_i, exp_i, X_a, Y_a, Ndiag, C_a) are handed over once, the
! first time a stage needs them.  The stage matrices stay in device memory
! between stages.  Their host globals (A_aij, B_ia, Q_ij, P_ia, P2_ia) are
! filled only where the reference reads or writes a stage file: a cache hit
! reads the file and uploads it, a file name other than 'None' downloads the
! stage and writes it.  A_aij and B_ia are always allocated, because the
! driver deallocates them after get_P (imcom.f90:159); L_i, K_a and T_ia are
! always filled on the host, as the reference fills them.  The lookup tables
! are built on the device and never allocated on the host.
!
! The reference's per-element helpers imcom_interp_lookup,
! imcom_A_plus_kNdiag and imcom_U/S_from_kappa(_test) have no counterpart
! here: the GPU kernels behind the stages above evaluate them on the device.
MODULE imcom_matrix

USE, INTRINSIC :: iso_c_binding
USE imcom_data
USE imcom_io
USE imcom_proc
USE imcom_gpu_iface

implicit none

! GPU state of the process: the C context, whether the tile's inputs are on
! the device, and the arithmetic of the Q^T B and Q W products (FP64 unless a
! driver asks for a split-precision scheme before the first stage).
type(c_ptr), save :: imcom_gpu_context = c_null_ptr
logical, save :: imcom_gpu_tile_ready = .false.
integer(c_int), save :: imcom_gpu_device = 0_c_int
integer(c_int), save :: imcom_gpu_precision_P = IMCOM_PREC_FP64
integer(c_int), save :: imcom_gpu_precision_T = IMCOM_PREC_FP64

CONTAINS

!---

SUBROUTINE imcom_gpu_open
! Opens the C context once per process.
implicit none
integer(c_int) :: st
if (c_associated(imcom_gpu_context)) return
st = imcom_gpu_init(imcom_gpu_context, imcom_gpu_device)
if (st.ne.0) then
  write(*, FMT='(A, I3)') "IMCOM ERROR: no usable GPU for the solve path, device ", imcom_gpu_device
  stop
end if
END SUBROUTINE imcom_gpu_open

!---

SUBROUTINE imcom_gpu_tile
! Hands the tile's input vectors to the device the first time a stage needs
! them (after imcom_build_Ndiag and imcom_calc_C, as the driver orders them).
implicit none
call imcom_gpu_open
if (imcom_gpu_tile_ready) return
if (.not.(allocated(x_i).and.allocated(y_i).and.allocated(exp_i).and.allocated(X_a) &
          .and.allocated(Y_a).and.allocated(Ndiag))) then
  write(*, FMT='(A)') "IMCOM ERROR: tile inputs (x_i, y_i, exp_i, X_a, Y_a, Ndiag) not built yet"
  stop
end if
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_set_tile(imcom_gpu_context, int(n, c_int64_t), &
                     int(m, c_int64_t), x_i, y_i, exp_i, X_a, Y_a, Ndiag, C_a), "set_tile")
! USERXY = 0 (imcom.f90:64-71): the driver built x_i, y_i, X_a, Y_a from the
! regular grids, whose pitches and dithers imcom_read_inconfigs allocated
! only in that case (imcom_io.f90:695-703); the lookups then interpolate each
! pixel separation once
if (allocated(gimxscale).and.allocated(gimyscale).and.allocated(dither) &
    .and.allocated(mask_subs_i)) then
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_set_grid(imcom_gpu_context, &
                       int(n1gim, c_int), int(n2gim, c_int), int(nexp, c_int), gimxscale, &
                       gimyscale, dither, mask_subs_i, int(n1out, c_int), int(n2out, c_int), &
                       outxscale, outyscale, outpos1, outpos2), "set_grid")
end if
imcom_gpu_tile_ready = .true.
END SUBROUTINE imcom_gpu_tile

!---

SUBROUTINE imcom_gpu_reset_tile
! Makes the next stage hand the tile's inputs over again (a driver that
! solves several tiles in one process calls this between tiles).
implicit none
imcom_gpu_tile_ready = .false.
END SUBROUTINE imcom_gpu_reset_tile

!---

SUBROUTINE imcom_read_stage(filename, what, n1_want, n2_want, arr)
! A cached stage file: its size must match the tile; the reference's
! cache-hit branches (imcom_matrix.f90:49-68 and alike) check the same.
implicit none
character(LEN=*), intent(IN) :: filename, what
integer, intent(IN) :: n1_want, n2_want
real(KIND=8), dimension(:, :), allocatable, intent(INOUT) :: arr
integer :: n1_tmp, n2_tmp, bit_tmp, alstat
call imcom_sizefits(trim(filename), n1_tmp, n2_tmp, bit_tmp)
if ((n1_tmp.ne.n1_want).or.(n2_tmp.ne.n2_want)) then
  write(*, FMT='(A)') "IMCOM ERROR: "//what//" in "//trim(filename)// &
                      " does not match the <config_file> image dimensions"
  write(*, FMT='(A, 2I9)') "IMCOM ERROR: expected ", n1_want, n2_want
  stop
end if
if (allocated(arr)) deallocate(arr)
allocate(arr(n1_tmp, n2_tmp), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for "//what
  stop
end if
write(*, FMT='(A)') "IMCOM: Reading "//what//" from "//trim(filename)
call imcom_readfits(trim(filename), n1_tmp, n2_tmp, arr)
END SUBROUTINE imcom_read_stage

!---

SUBROUTINE imcom_host_matrix(arr, n1_want, n2_want, what)
! Allocates a host global for a stage (its contents are filled only when the
! stage is downloaded).
implicit none
real(KIND=8), dimension(:, :), allocatable, intent(INOUT) :: arr
integer, intent(IN) :: n1_want, n2_want
character(LEN=*), intent(IN) :: what
integer :: alstat
if (allocated(arr)) deallocate(arr)
allocate(arr(n1_want, n2_want), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for "//what//" -- too large?"
  stop
end if
END SUBROUTINE imcom_host_matrix

!---

SUBROUTINE imcom_get_A(npoly, npad, saveA, forcebuild)
! A from AFILE when it exists and no rebuild is forced, else built on the device.
implicit none
integer, intent(IN) :: npoly, npad, saveA, forcebuild
logical :: Aexists
inquire(FILE=trim(Afile), EXIST=Aexists)
if (Aexists.and.(forcebuild.eq.0)) then
  call imcom_read_stage(Afile, "A matrix", n, n, A_aij)
  call imcom_gpu_tile
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_set_A(imcom_gpu_context, A_aij), "set_A")
else
  call imcom_build_Alookup(npad)
  call imcom_lookup_A(npoly, npad, saveA)
end if
END SUBROUTINE imcom_get_A

!---

SUBROUTINE imcom_get_B(npoly, npad, saveB, forcebuild)
! B from BFILE when it exists and no rebuild is forced, else built on the device.
implicit none
integer, intent(IN) :: npoly, npad, saveB, forcebuild
logical :: Bexists
inquire(FILE=trim(Bfile), EXIST=Bexists)
if (Bexists.and.(forcebuild.eq.0)) then
  call imcom_read_stage(Bfile, "B matrix", n, m, B_ia)
  call imcom_gpu_tile
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_set_B(imcom_gpu_context, B_ia), "set_B")
else
  call imcom_build_Blookup(npad)
  call imcom_lookup_B(npoly, npad, saveB)
end if
END SUBROUTINE imcom_get_B

!---

SUBROUTINE imcom_get_QL(saveQL, forcebuild)
! Q and L from QFILE / LFILE when both exist and no rebuild is forced, else
! the eigendecomposition of the device's A.
implicit none
integer, intent(IN) :: saveQL, forcebuild
logical :: Qexists, Lexists
real(KIND=8), dimension(:, :), allocatable :: Ltmp
integer :: alstat
inquire(FILE=trim(Qfile), EXIST=Qexists)
inquire(FILE=trim(Lfile), EXIST=Lexists)
if (Qexists.and.Lexists.and.(forcebuild.eq.0)) then
  call imcom_read_stage(Qfile, "Q matrix", n, n, Q_ij)
  call imcom_read_stage(Lfile, "L eigenvalues", n, 1, Ltmp)
  if (allocated(L_i)) deallocate(L_i)
  allocate(L_i(n), STAT=alstat)
  if (alstat.ne.0) then
    write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for L vector"
    stop
  end if
  L_i = Ltmp(:, 1)
  deallocate(Ltmp)
  call imcom_gpu_tile
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_set_QL(imcom_gpu_context, Q_ij, L_i), "set_QL")
else
  call imcom_eigen_A(saveQL)
end if
END SUBROUTINE imcom_get_QL

!---

SUBROUTINE imcom_get_P(saveP, forcebuild)
! P from PFILE when it exists and no rebuild is forced (P2 = P*P on the host,
! as the reference keeps it), else P = Q^T B on the device.
implicit none
integer, intent(IN) :: saveP, forcebuild
logical :: Pexists
inquire(FILE=trim(Pfile), EXIST=Pexists)
if (Pexists.and.(forcebuild.eq.0)) then
  call imcom_read_stage(Pfile, "P matrix", n, m, P_ia)
  call imcom_host_matrix(P2_ia, n, m, "P^2 matrix")
  P2_ia = P_ia * P_ia
  call imcom_gpu_tile
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_set_P(imcom_gpu_context, P_ia), "set_P")
else
  call imcom_build_P(saveP)
end if
END SUBROUTINE imcom_get_P

!---

SUBROUTINE imcom_build_Alookup(npad)
! The A lookup tables from the PSF spectra Gt_rot, on the device (they stay
! there for imcom_lookup_A).
implicit none
integer, intent(IN) :: npad
call imcom_gpu_open
write(*, FMT='(A)') "IMCOM: Building A lookup tables on the GPU"
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_build_Alookup(imcom_gpu_context, &
                     int(psfconst, c_int), int(nexppsf, c_int), int(n1psf, c_int), &
                     int(n2psf, c_int), int(npad, c_int), Gt_rot, c_null_ptr), "build_Alookup")
END SUBROUTINE imcom_build_Alookup

!---

SUBROUTINE imcom_build_Blookup(npad)
! The B lookup tables from Gammat and Gt_rot, on the device.
implicit none
integer, intent(IN) :: npad
call imcom_gpu_open
write(*, FMT='(A)') "IMCOM: Building B lookup tables on the GPU"
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_build_Blookup(imcom_gpu_context, &
                     int(psfconst, c_int), int(nexppsf, c_int), int(n1psf, c_int), &
                     int(n2psf, c_int), int(npad, c_int), Gammat, Gt_rot, c_null_ptr), &
                     "build_Blookup")
END SUBROUTINE imcom_build_Blookup

!---

SUBROUTINE imcom_calc_C
! C = sum |Gammat|^2 / (n1psf n2psf) on the device.
implicit none
call imcom_gpu_open
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_calc_C(imcom_gpu_context, int(n1psf, c_int), &
                     int(n2psf, c_int), Gammat, C_a), "calc_C")
write(*, FMT='(A, E24.16)') "IMCOM: C_a = ", C_a
END SUBROUTINE imcom_calc_C

!---

SUBROUTINE imcom_build_Ndiag
! The noise diagonal, host set-up kept as the reference defines it: the n
! pixels are cut into nexp equal runs of n / nexp (the last run takes the
! remainder) and run e gets noise(e), whatever the mask left of exposure e.
implicit none
integer :: i, e, run, alstat
if (allocated(Ndiag)) deallocate(Ndiag)
allocate(Ndiag(n), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate space for noise variance vector"
  stop
end if
run = n / nexp
do i = 1, n
  if (run.gt.0) then
    e = min((i - 1) / run + 1, nexp)
  else
    e = nexp
  end if
  Ndiag(i) = noise(e)
end do
END SUBROUTINE imcom_build_Ndiag

!---

SUBROUTINE imcom_lookup_A(npoly, npad, saveA)
! A on the device from its lookup tables; downloaded and written to AFILE
! when saveA = 1.
implicit none
integer, intent(IN) :: npoly, npad, saveA
type(c_ptr) :: out
call imcom_gpu_tile
write(*, FMT='(A)') "IMCOM: Building A matrix from lookup table on the GPU"
call imcom_host_matrix(A_aij, n, n, "A matrix")
out = c_null_ptr
if (saveA.eq.1) out = imcom_gpu_addr(A_aij)
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_lookup_A(imcom_gpu_context, int(npoly, c_int), &
                     int(npad, c_int), psfxscale, psfyscale, out), "lookup_A")
if (saveA.eq.1) then
  write(*, FMT='(A)') "IMCOM: Saving A matrix to "//trim(Afile)
  call imcom_writefits(trim(Afile), n, n, A_aij)
end if
END SUBROUTINE imcom_lookup_A

!---

SUBROUTINE imcom_lookup_B(npoly, npad, saveB)
! B on the device from its lookup tables; downloaded and written to BFILE
! when saveB = 1.
implicit none
integer, intent(IN) :: npoly, npad, saveB
type(c_ptr) :: out
call imcom_gpu_tile
write(*, FMT='(A)') "IMCOM: Building B vectors from lookup table on the GPU"
call imcom_host_matrix(B_ia, n, m, "B matrix")
out = c_null_ptr
if (saveB.eq.1) out = imcom_gpu_addr(B_ia)
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_lookup_B(imcom_gpu_context, int(npoly, c_int), &
                     int(npad, c_int), psfxscale, psfyscale, out), "lookup_B")
if (saveB.eq.1) then
  write(*, FMT='(A)') "IMCOM: Saving B vectors to "//trim(Bfile)
  call imcom_writefits(trim(Bfile), n, m, B_ia)
end if
END SUBROUTINE imcom_lookup_B

!---

SUBROUTINE imcom_eigen_A(saveQL)
! Eigendecomposition of the device's A: Q, L in descending eigenvalue order
! and the condition number.  L comes back to the host; Q only when it is
! written to QFILE (saveQL /= 0).
implicit none
integer, intent(IN) :: saveQL
type(c_ptr) :: qout
integer :: alstat
call imcom_gpu_tile
if (allocated(L_i)) deallocate(L_i)
allocate(L_i(n), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for Eigendecomposition outputs"
  stop
end if
qout = c_null_ptr
if (saveQL.ne.0) then
  call imcom_host_matrix(Q_ij, n, n, "Q matrix")
  qout = imcom_gpu_addr(Q_ij)
end if
write(*, FMT='(A)') "IMCOM: Performing eigendecomposition of A matrix on the GPU"
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_eigen_A(imcom_gpu_context, qout, &
                     imcom_gpu_addr(L_i), condition), "eigen_A")
if (condition.lt.0.d0) then
  write(*, FMT='(A)') "IMCOM WARNING: Negative eigenvalues!"
  write(*, FMT='(A, E16.8)') "IMCOM WARNING: L_max = ", maxval(L_i)
  write(*, FMT='(A, E16.8)') "IMCOM WARNING: L_min = ", minval(L_i)
else
  write(*, FMT='(A, E24.16)') "IMCOM: A matrix condition number = ", condition
end if
if (saveQL.ne.0) then
  write(*, FMT='(A)') "IMCOM: Writing eigenvectors to "//trim(Qfile)
  call imcom_writefits(trim(Qfile), n, n, Q_ij)
  write(*, FMT='(A)') "IMCOM: Writing eigenvalues to "//trim(Lfile)
  call imcom_writefits(trim(Lfile), n, 1, L_i)
end if
END SUBROUTINE imcom_eigen_A

!---

SUBROUTINE imcom_build_P(saveP)
! P = Q^T B on the device; downloaded (with P2 = P*P) and written to PFILE
! when saveP = 1.
implicit none
integer, intent(IN) :: saveP
type(c_ptr) :: out
write(*, FMT='(A)') "IMCOM: Building projection matrix on the GPU"
out = c_null_ptr
if (saveP.eq.1) then
  call imcom_host_matrix(P_ia, n, m, "P matrix")
  out = imcom_gpu_addr(P_ia)
end if
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_build_P(imcom_gpu_context, &
                     imcom_gpu_precision_P, out), "build_P")
if (saveP.eq.1) then
  write(*, FMT='(A)') "IMCOM: Writing P matrix to "//trim(Pfile)
  call imcom_writefits(trim(Pfile), n, m, P_ia)
  call imcom_host_matrix(P2_ia, n, m, "P^2 matrix")
  P2_ia = P_ia * P_ia
end if
END SUBROUTINE imcom_build_P

!---

SUBROUTINE imcom_build_T
! T = -1/2 Q (P / (L + K_a)) on the device, downloaded and always written to
! TFILE, as the reference does.
implicit none
write(*, FMT='(A)') "IMCOM: Building T matrix on the GPU"
call imcom_host_matrix(T_ia, n, m, "T matrix")
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_build_T(imcom_gpu_context, &
                     imcom_gpu_precision_T, imcom_gpu_addr(T_ia)), "build_T")
write(*, FMT='(A)') "IMCOM: Writing T matrix to "//trim(Tfile)
call imcom_writefits(trim(Tfile), n, m, T_ia)
END SUBROUTINE imcom_build_T

!---

SUBROUTINE imcom_build_image
! H = T^T I on the device, reshaped to the output grid and written to HFILE;
! I_i is released afterwards, as the reference does.
implicit none
integer :: alstat
call imcom_gpu_tile
write(*, FMT='(A)') "IMCOM: Building output image H"
if (allocated(H_a)) deallocate(H_a)
allocate(H_a(m), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate space for image output vectors"
  stop
end if
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_build_image(imcom_gpu_context, I_i, H_a), &
                     "build_image")
deallocate(I_i, STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot deallocate space for image output vector I_i"
  stop
end if
if (allocated(H)) deallocate(H)
allocate(H(n1out, n2out), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate space for output image matrix H"
  stop
end if
H = reshape(H_a, (/ n1out, n2out /))
deallocate(H_a)
write(*, FMT='(A)') "IMCOM: Saving H vector to "//trim(Hfile)
call imcom_writefits(trim(Hfile), n1out, n2out, H)
END SUBROUTINE imcom_build_image

!---

SUBROUTINE imcom_build_U
! The leakage map U(K_a) on the device (noise-target mode), written to UFILE.
implicit none
integer :: alstat
write(*, FMT='(A)') "IMCOM: Building output leakage map U"
if (allocated(U_a)) deallocate(U_a)
allocate(U_a(m), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for leakage map U"
  stop
end if
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_build_U(imcom_gpu_context, U_a), "build_U")
if (allocated(U)) deallocate(U)
allocate(U(n1out, n2out), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for output leakage matrix U"
  stop
end if
U = reshape(U_a, (/ n1out, n2out /))
deallocate(U_a)
write(*, FMT='(A)') "IMCOM: Saving U vector to "//trim(Ufile)
call imcom_writefits(trim(Ufile), n1out, n2out, U)
END SUBROUTINE imcom_build_U

!---

SUBROUTINE imcom_build_S
! The noise map S = sum_i T_ia^2 Ndiag(i) on the device, written to SFILE.
implicit none
integer :: alstat
write(*, FMT='(A)') "IMCOM: Building output noise map S"
if (allocated(S_a)) deallocate(S_a)
allocate(S_a(m), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate space for noise map output"
  stop
end if
call imcom_gpu_check(imcom_gpu_context, imcom_gpu_build_S(imcom_gpu_context, S_a), "build_S")
if (allocated(S)) deallocate(S)
allocate(S(n1out, n2out), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate space for noise map output matrix"
  stop
end if
S = reshape(S_a, (/ n1out, n2out /))
deallocate(S_a)
write(*, FMT='(A)') "IMCOM: Saving S vector to "//trim(Sfile)
call imcom_writefits(trim(Sfile), n1out, n2out, S)
END SUBROUTINE imcom_build_S

!---

END MODULE imcom_matrix
