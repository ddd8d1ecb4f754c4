! ISO_C_BINDING interface of libportal_hip.so (include/portal_hip.h) for PORTAL's
! Fortran drivers.  Arrays are passed as the drivers hold them (column-major,
! 1-based ids); every function returns a status code (PORTAL_OK = 0) where the
! reference would `stop`; portal_last_error gives the message.
module portal_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PORTAL_OK = 0, PORTAL_ERR_ARG = -1, PORTAL_ERR_DEVICE = -2, &
       PORTAL_ERR_DEPTH = -3, PORTAL_ERR_DEPTH_RT = -4, PORTAL_ERR_LU = -5, PORTAL_ERR_TEMP = -6, &
       PORTAL_ERR_STATE = -7, PORTAL_ERR_NOMEM = -8, PORTAL_ERR_UNSUPPORTED = -9
  integer(c_int), parameter :: PORTAL_FP64 = 0, PORTAL_MIXED = 1

  type, bind(C) :: portal_se_params
    integer(c_int) :: n_sphere
    type(c_ptr) :: xth
    integer(c_int) :: nvel
    real(c_double) :: dx
    type(c_ptr) :: gvec
    type(c_ptr) :: bvec_all
    integer(c_int) :: n_depth
    integer(c_int) :: extra_info
    integer(c_int) :: precision
    integer(c_int) :: star_mod
    integer(c_int) :: nth, nph
    real(c_double) :: R_star, T_star
    type(c_ptr) :: X_star
  end type

  type, bind(C) :: portal_iter_params
    integer(c_int) :: max_iter
    real(c_double) :: tol_pop
    real(c_double) :: tol_align
    integer(c_int) :: surrogate
    real(c_double) :: fallback_tol
    integer(c_int) :: rebuild_paths
    integer(c_int) :: ng_accel
    integer(c_int) :: ali
  end type

  type, bind(C) :: portal_img_params
    integer(c_int) :: nlevels
    type(c_ptr) :: l1, l2, nu0_all, Aji_all, kd_all
    real(c_double) :: th_inc, ph_az
    integer(c_int) :: nvel
    type(c_ptr) :: vel_out
    integer(c_int) :: im_pix
    real(c_double) :: im_size, dx, alpha
    integer(c_int) :: n_depth_rt, precision
    integer(c_int) :: star_mod
    real(c_double) :: R_star, T_star
    type(c_ptr) :: X_star
    integer(c_int) :: node_pixels
  end type

  type, bind(C) :: portal_stats
    real(c_double) :: t_walk_ms, t_prep_ms, t_rt_ms, t_se_ms, t_img_ms, t_total_ms, rt_kernel_avg_ms
    integer(c_long_long) :: n_cross, n_paths
    integer(c_int) :: iterations, fallbacks, lu_failures
    real(c_double) :: max_rel_change
    integer(c_int) :: rt_variant, rt_waves, max_path, ng_steps, converged
    real(c_double) :: max_align_change
    integer(c_int) :: se_form, se_waves
    real(c_double) :: t_mlp_ms
    integer(c_int) :: rt_order
  end type

  interface
    integer(c_int) function portal_ctx_create(ctx, device) bind(C, name="portal_ctx_create")
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function

    integer(c_int) function portal_ctx_destroy(ctx) bind(C, name="portal_ctx_destroy")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function portal_last_error_c(ctx) bind(C, name="portal_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function portal_upload_grid(ctx, node_num, bulk_node_num, XYZ, vel_all, temp_all, b_all, &
        h2_dens, ab_all, dust_p, del_ne, del_num_ne, maxne, tx) bind(C, name="portal_upload_grid")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: node_num, bulk_node_num, maxne
      real(c_double), intent(in) :: XYZ(*), vel_all(*), temp_all(*), b_all(*), h2_dens(*), ab_all(*), dust_p(*)
      integer(c_int), intent(in) :: del_ne(*), del_num_ne(*)
      real(c_double), value :: tx
    end function

    integer(c_int) function portal_upload_species(ctx, e_num, j_lev, e_levs_hz, r_num, r_trans, t_Aji, nu_hz, &
        t_Bij, kd_v, c_num, c_trans, c_temp, c_par, t_Cji, col_temps, kmax) bind(C, name="portal_upload_species")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: e_num, r_num, c_num, c_temp, c_par, kmax
      integer(c_int), intent(in) :: j_lev(*), r_trans(*), c_trans(*)
      real(c_double), intent(in) :: e_levs_hz(*), t_Aji(*), nu_hz(*), t_Bij(*), kd_v(*), t_Cji(*), col_temps(*)
    end function

    integer(c_int) function portal_set_levpop(ctx, lev_p) bind(C, name="portal_set_levpop")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: lev_p(*)
    end function

    ! collisional depolarisation (depol.f90:242-359): the table of de_polarization,
    ! D0_all(nT, nlev, 2:kmax), and switching it on (c_null_ptr: off, as main_sub.f90:96)
    integer(c_int) function portal_depol_table(ctx, D0_all) bind(C, name="portal_depol_table")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: D0_all(*)
    end function

    integer(c_int) function portal_set_depol(ctx, D0_all) bind(C, name="portal_set_depol")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(c_ptr), value :: D0_all
    end function

    integer(c_int) function portal_get_levpop(ctx, lev_p) bind(C, name="portal_get_levpop")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: lev_p(*)
    end function

    integer(c_int) function portal_se_pass(ctx, p, pol_pops, pol_rates, JJ_20, stats) bind(C, name="portal_se_pass")
      import :: c_ptr, c_int, c_double, portal_se_params, portal_stats
      type(c_ptr), value :: ctx
      type(portal_se_params), intent(in) :: p
      real(c_double), intent(out) :: pol_pops(*), pol_rates(*), JJ_20(*)
      type(portal_stats), intent(out) :: stats
    end function

    integer(c_int) function portal_lambda_iterate(ctx, p, it, pol_pops, lev_p_out, stats) &
        bind(C, name="portal_lambda_iterate")
      import :: c_ptr, c_int, c_double, portal_se_params, portal_iter_params, portal_stats
      type(c_ptr), value :: ctx
      type(portal_se_params), intent(in) :: p
      type(portal_iter_params), intent(in) :: it
      real(c_double), intent(out) :: pol_pops(*), lev_p_out(*)
      type(portal_stats), intent(out) :: stats
    end function

    integer(c_int) function portal_image_pass(ctx, p, pol_pops, bvec_all, I, Q, U, TAU, dust, stats) &
        bind(C, name="portal_image_pass")
      import :: c_ptr, c_int, c_double, portal_img_params, portal_stats
      type(c_ptr), value :: ctx
      type(portal_img_params), intent(in) :: p
      real(c_double), intent(in) :: pol_pops(*), bvec_all(*)
      real(c_double), intent(out) :: I(*), Q(*), U(*), TAU(*)
      type(c_ptr), value :: dust
      type(portal_stats), intent(out) :: stats
    end function
  end interface

contains

  function portal_last_error(ctx) result(msg)
    type(c_ptr), value :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: k
    msg = ' '
    call c_f_pointer(portal_last_error_c(ctx), s, [256])
    do k = 1, 256
      if (s(k) == c_null_char) exit
      msg(k:k) = s(k)
    end do
  end function

end module portal_hip
