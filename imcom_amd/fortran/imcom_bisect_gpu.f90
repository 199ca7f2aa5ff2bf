! MODULE imcom_bisect for the MI355X solve path: a drop-in for the
! reference's imcom_bisect.f90, compiled after imcom_matrix_gpu.f90.  Public
! procedures kept with their names and arguments:
!
!   imcom_get_T(forcebuild)                                  imcom_bisect.f90:40-74
!   imcom_bisect_kappa(UnotS, US, k_min, k_max, tol, maxnits) imcom_bisect.f90:78-170
!
! imcom_get_T reads TFILE instead of solving when it exists and forcebuild is
! 0, as the reference does, and hands T to the device for imcom_build_image.
! imcom_bisect_kappa runs every output pixel's bisection on the GPU, one
! pixel per workgroup, with the reference's control flow (imcom_bisect_by_U /
! imcom_bisect_by_S, imcom_bisect.f90:174-267, over imcom_U/S_from_kappa);
! those per-pixel routines have no separate counterpart here.  With
! imcom_gpu_learned set (and the network parameters uploaded through
! imcom_gpu_mlp_set_params), each pixel's bisection starts from the learned
! kappa bracket instead, falling back to the full range where the bracket
! does not hold the root; the default is the reference's full range.
MODULE imcom_bisect

USE, INTRINSIC :: iso_c_binding
USE imcom_data
USE imcom_io
USE imcom_proc
USE imcom_matrix
USE imcom_gpu_iface

implicit none

logical, save :: imcom_gpu_learned = .false.
real(c_float), save :: imcom_gpu_zsig = 3.0_c_float

CONTAINS

!---

SUBROUTINE imcom_get_T(forcebuild)
! T from TFILE when it exists and no rebuild is forced, else kappa by
! bisection, T, and the S (leakage mode) or U (noise mode) map.
implicit none
integer, intent(IN) :: forcebuild
logical :: Texists
inquire(FILE=trim(Tfile), EXIST=Texists)
if (Texists.and.(forcebuild.eq.0)) then
  call imcom_read_stage(Tfile, "T matrix", n, m, T_ia)
  call imcom_gpu_tile
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_set_T(imcom_gpu_context, T_ia), "set_T")
else
  call imcom_bisect_kappa(USB, US_max, kappa_min, kappa_max, US_tol, maxNbis)
  call imcom_build_T
  if (USB) then
    call imcom_build_S
  else
    call imcom_build_U
  end if
end if
END SUBROUTINE imcom_get_T

!---

SUBROUTINE imcom_bisect_kappa(UnotS, US, k_min, k_max, tol, maxnits)
! kappa per output pixel meeting the leakage (UnotS) or noise target US; K_a
! and U_a or S_a on the host, the U or S map and the K map written to UFILE /
! SFILE and KFILE.
implicit none
logical, intent(IN) :: UnotS
real(KIND=8), intent(IN) :: US, k_min, k_max, tol
integer, intent(IN) :: maxnits
real(KIND=8), dimension(:), allocatable :: USa
integer :: alstat
integer(c_int) :: mode
call imcom_gpu_tile
write(*, FMT='(A)') "IMCOM: Solving for kappa using interval bisection on the GPU"
if (allocated(K_a)) deallocate(K_a)
allocate(K_a(m), USa(m), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for kappa and U/S storage arrays"
  stop
end if
mode = 0_c_int
if (UnotS) mode = 1_c_int
if (imcom_gpu_learned) then
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_bisect_kappa_learned(imcom_gpu_context, mode, &
                       US, k_min, k_max, tol, int(maxnits, c_int), imcom_gpu_zsig, K_a, USa, &
                       c_null_ptr, c_null_ptr), "bisect_kappa_learned")
else
  call imcom_gpu_check(imcom_gpu_context, imcom_gpu_bisect_kappa(imcom_gpu_context, mode, US, &
                       k_min, k_max, tol, int(maxnits, c_int), c_null_ptr, K_a, USa, &
                       c_null_ptr), "bisect_kappa")
end if
write(*, FMT='(A)') "IMCOM: 100% complete"
if (UnotS) then
  if (allocated(U)) deallocate(U)
  allocate(U(n1out, n2out), STAT=alstat)
  if (alstat.ne.0) then
    write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for U output array"
    stop
  end if
  U = reshape(USa, (/ n1out, n2out /))
  write(*, FMT='(A)') "IMCOM: Saving U leakage map to "//trim(Ufile)
  call imcom_writefits(trim(Ufile), n1out, n2out, U)
else
  if (allocated(S)) deallocate(S)
  allocate(S(n1out, n2out), STAT=alstat)
  if (alstat.ne.0) then
    write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for S output array"
    stop
  end if
  S = reshape(USa, (/ n1out, n2out /))
  write(*, FMT='(A)') "IMCOM: Saving S variance map to "//trim(Sfile)
  call imcom_writefits(trim(Sfile), n1out, n2out, S)
end if
deallocate(USa)
if (allocated(K)) deallocate(K)
allocate(K(n1out, n2out), STAT=alstat)
if (alstat.ne.0) then
  write(*, FMT='(A)') "IMCOM ERROR: Cannot allocate memory for kappa output array"
  stop
end if
K = reshape(K_a, (/ n1out, n2out /))
write(*, FMT='(A)') "IMCOM: Saving K solution map to "//trim(Kfile)
call imcom_writefits(trim(Kfile), n1out, n2out, K)
END SUBROUTINE imcom_bisect_kappa

!---

END MODULE imcom_bisect
