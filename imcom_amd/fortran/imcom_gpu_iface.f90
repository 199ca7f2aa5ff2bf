! ISO_C_BINDING interface to libimcom_hip.so (include/imcom_gpu.h) for the
! reference's Fortran driver.  The replacement imcom_matrix / imcom_bisect
! modules (imcom_matrix_gpu.f90, imcom_bisect_gpu.f90) call these from the
! bodies of imcom_get_A/B/QL/P, imcom_lookup_A/B, imcom_eigen_A, imcom_build_P,
! imcom_get_T, imcom_bisect_kappa, imcom_build_T/S/U and imcom_build_image,
! keeping every public procedure name and argument, so imcom.f90 compiles
! unchanged against them.  Sizes are int64, inputs are assumed-size
! real(c_double) (Fortran column-major, as the C side expects).  A stage
! output that may stay on the device is a type(c_ptr), value argument: pass
! imcom_gpu_addr(array) to receive it or c_null_ptr to leave it on the
! device.  A C status other than 0 is reported as "IMCOM ERROR: <message>"
! followed by STOP, as the reference does.
MODULE imcom_gpu_iface
  USE, INTRINSIC :: iso_c_binding
  implicit none

  integer(c_int), parameter :: IMCOM_PREC_FP64 = 0, IMCOM_PREC_FP32X3 = 1, &
                               IMCOM_PREC_BF16X6 = 2
  integer(c_int), parameter :: IMCOM_FETCH_A = 0, IMCOM_FETCH_Q = 1, IMCOM_FETCH_L = 2, &
                               IMCOM_FETCH_B = 3, IMCOM_FETCH_P = 4, IMCOM_FETCH_T = 5

  interface
    integer(c_int) function imcom_gpu_init(ctx, device) bind(C, name="imcom_gpu_init")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function
    subroutine imcom_gpu_finalize(ctx) bind(C, name="imcom_gpu_finalize")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function imcom_gpu_last_error(ctx) bind(C, name="imcom_gpu_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function imcom_gpu_set_tile(ctx, n, m, x_i, y_i, exp_i, X_a, Y_a, Ndiag, C_a) &
        bind(C, name="imcom_gpu_set_tile")
      import :: c_int, c_ptr, c_int64_t, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n, m
      real(c_double), intent(in) :: x_i(*), y_i(*), X_a(*), Y_a(*), Ndiag(*)
      integer(c_int32_t), intent(in) :: exp_i(*)
      real(c_double), value :: C_a
    end function
    integer(c_int) function imcom_gpu_set_grid(ctx, n1gim, n2gim, nexp, gimxscale, gimyscale, &
                                               dither, mask_subs_i, n1out, n2out, outxscale, &
                                               outyscale, outpos1, outpos2) &
        bind(C, name="imcom_gpu_set_grid")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n1gim, n2gim, nexp, n1out, n2out
      real(c_double), intent(in) :: gimxscale(*), gimyscale(*), dither(*)
      integer(c_int32_t), intent(in) :: mask_subs_i(*)
      real(c_double), value :: outxscale, outyscale, outpos1, outpos2
    end function
    integer(c_int) function imcom_gpu_set_tables(ctx, psfconst, nexppsf, n1, n2, Alookup, Blookup) &
        bind(C, name="imcom_gpu_set_tables")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: psfconst, nexppsf, n1, n2
      real(c_double), intent(in) :: Alookup(*), Blookup(*)
    end function
    ! device lookup-table set-up: spectra are complex(c_double_complex),
    ! whose storage is the interleaved (re, im) layout the C side reads
    integer(c_int) function imcom_gpu_assemble_inputs(ctx, n1gim, n2gim, nexp, Im_unmasked, &
        saturation, gimxscale, gimyscale, dither, n_out, x_out, y_out, exp_out, I_out) &
        bind(C, name="imcom_gpu_assemble_inputs")
      import :: c_int, c_ptr, c_double, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: n1gim, n2gim, nexp
      real(c_double), intent(in) :: Im_unmasked(*), gimxscale(*), gimyscale(*), dither(*)
      real(c_double), value :: saturation
      integer(c_int64_t), intent(out) :: n_out
      real(c_double), intent(out) :: x_out(*), y_out(*), I_out(*)
      integer(c_int32_t), intent(out) :: exp_out(*)
    end function
    integer(c_int) function imcom_gpu_rotate_image(ctx, n1, n2, im_in, theta_deg, npoly, &
        conserve_flux, im_out) bind(C, name="imcom_gpu_rotate_image")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n1, n2, npoly, conserve_flux
      real(c_double), intent(in) :: im_in(*)
      real(c_double), value :: theta_deg
      real(c_double), intent(out) :: im_out(*)
    end function
    integer(c_int) function imcom_gpu_ft_r2c(ctx, n1, n2, count, img, spec_out) &
        bind(C, name="imcom_gpu_ft_r2c")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: n1, n2, count
      real(c_double), intent(in) :: img(*)
      complex(c_double_complex), intent(out) :: spec_out(*)
    end function
    integer(c_int) function imcom_gpu_build_Alookup(ctx, psfconst, nexppsf, n1psf, n2psf, npad, &
        Gt_rot, Alookup_out) bind(C, name="imcom_gpu_build_Alookup")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: psfconst, nexppsf, n1psf, n2psf, npad
      complex(c_double_complex), intent(in) :: Gt_rot(*)
      type(c_ptr), value :: Alookup_out
    end function
    integer(c_int) function imcom_gpu_build_Blookup(ctx, psfconst, nexppsf, n1psf, n2psf, npad, &
        Gammat, Gt_rot, Blookup_out) bind(C, name="imcom_gpu_build_Blookup")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: psfconst, nexppsf, n1psf, n2psf, npad
      complex(c_double_complex), intent(in) :: Gammat(*), Gt_rot(*)
      type(c_ptr), value :: Blookup_out
    end function
    integer(c_int) function imcom_gpu_calc_C(ctx, n1psf, n2psf, Gammat, C_out) &
        bind(C, name="imcom_gpu_calc_C")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: n1psf, n2psf
      complex(c_double_complex), intent(in) :: Gammat(*)
      real(c_double), intent(out) :: C_out
    end function
    integer(c_int) function imcom_gpu_lookup_A(ctx, npoly, npad, psfxscale, psfyscale, A_out) &
        bind(C, name="imcom_gpu_lookup_A")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npoly, npad
      real(c_double), value :: psfxscale, psfyscale
      type(c_ptr), value :: A_out
    end function
    integer(c_int) function imcom_gpu_set_A(ctx, A) bind(C, name="imcom_gpu_set_A")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: A(*)
    end function
    integer(c_int) function imcom_gpu_eigen_A(ctx, Q_out, L_out, condition) &
        bind(C, name="imcom_gpu_eigen_A")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      type(c_ptr), value :: Q_out, L_out
      real(c_double), intent(out) :: condition
    end function
    integer(c_int) function imcom_gpu_set_QL(ctx, Q, L) bind(C, name="imcom_gpu_set_QL")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: Q(*), L(*)
    end function
    integer(c_int) function imcom_gpu_lookup_B(ctx, npoly, npad, psfxscale, psfyscale, B_out) &
        bind(C, name="imcom_gpu_lookup_B")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: npoly, npad
      real(c_double), value :: psfxscale, psfyscale
      type(c_ptr), value :: B_out
    end function
    integer(c_int) function imcom_gpu_set_B(ctx, B) bind(C, name="imcom_gpu_set_B")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: B(*)
    end function
    integer(c_int) function imcom_gpu_build_P(ctx, precision, P_out) bind(C, name="imcom_gpu_build_P")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      type(c_ptr), value :: P_out
    end function
    integer(c_int) function imcom_gpu_set_P(ctx, P) bind(C, name="imcom_gpu_set_P")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: P(*)
    end function
    integer(c_int) function imcom_gpu_bisect_kappa(ctx, UnotS, US_max, k_min, k_max, tol, maxnits, &
        bracket, K_out, US_out, evals_out) bind(C, name="imcom_gpu_bisect_kappa")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: UnotS, maxnits
      real(c_double), value :: US_max, k_min, k_max, tol
      type(c_ptr), value :: bracket
      real(c_double), intent(out) :: K_out(*), US_out(*)
      type(c_ptr), value :: evals_out
    end function
    integer(c_int) function imcom_gpu_bisect_kappa_learned(ctx, UnotS, US_max, k_min, k_max, tol, &
        maxnits, zsig, K_out, US_out, evals_out, hits_out) &
        bind(C, name="imcom_gpu_bisect_kappa_learned")
      import :: c_int, c_ptr, c_double, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: UnotS, maxnits
      real(c_double), value :: US_max, k_min, k_max, tol
      real(c_float), value :: zsig
      real(c_double), intent(out) :: K_out(*), US_out(*)
      type(c_ptr), value :: evals_out, hits_out
    end function
    integer(c_int) function imcom_gpu_mlp_set_params(ctx, params) &
        bind(C, name="imcom_gpu_mlp_set_params")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: params(*)
    end function
    integer(c_int) function imcom_gpu_build_T(ctx, precision, T_out) bind(C, name="imcom_gpu_build_T")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
      type(c_ptr), value :: T_out
    end function
    integer(c_int) function imcom_gpu_set_T(ctx, T) bind(C, name="imcom_gpu_set_T")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: T(*)
    end function
    integer(c_int) function imcom_gpu_build_S(ctx, S_out) bind(C, name="imcom_gpu_build_S")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: S_out(*)
    end function
    integer(c_int) function imcom_gpu_build_U(ctx, U_out) bind(C, name="imcom_gpu_build_U")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: U_out(*)
    end function
    integer(c_int) function imcom_gpu_build_image(ctx, I_i, H_out) bind(C, name="imcom_gpu_build_image")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: I_i(*)
      real(c_double), intent(out) :: H_out(*)
    end function
    ! stage files copied out while the next stages run: the host array must
    ! stay allocated and untouched until imcom_gpu_fetch_wait returns
    integer(c_int) function imcom_gpu_fetch_async(ctx, what, host) bind(C, name="imcom_gpu_fetch_async")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: what
      real(c_double) :: host(*)
    end function
    integer(c_int) function imcom_gpu_fetch_wait(ctx, what) bind(C, name="imcom_gpu_fetch_wait")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: what
    end function
  end interface

CONTAINS

  ! Address of a contiguous real(c_double) array, for the type(c_ptr) stage
  ! outputs above (the array itself need not carry TARGET).
  FUNCTION imcom_gpu_addr(a) RESULT(p)
    real(c_double), target, intent(inout) :: a(*)
    type(c_ptr) :: p
    p = c_loc(a(1))
  END FUNCTION imcom_gpu_addr

  ! Reference-style error stop: print "IMCOM ERROR: <library message>" and stop.
  SUBROUTINE imcom_gpu_check(ctx, status, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    interface
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
      end function
    end interface
    type(c_ptr) :: p
    integer(c_size_t) :: len
    if (status.eq.0) return
    p = imcom_gpu_last_error(ctx)
    len = c_strlen(p)
    call c_f_pointer(p, msg, [len])
    write(*, FMT='(A)', ADVANCE='NO') "IMCOM ERROR: "//where//": "
    do i = 1, int(len)
      write(*, FMT='(A)', ADVANCE='NO') msg(i)
    end do
    write(*, *)
    stop
  END SUBROUTINE imcom_gpu_check

END MODULE imcom_gpu_iface
