! Fortran caller of the C ABI: the node loop of an SE pass replaced by one call,
! as a PORTAL driver would make it.  Reads a model dumped by the Python tests
! (stream file: sizes, then the driver's arrays in the reference's layouts),
! runs portal_se_pass and writes pol_pops (node_num x 2L) to a stream file and,
! given a third argument, to a text file with the reference's own statements
! (main_para.f90:271-275: one list-directed record per node), the file that the
! image-pass driver reads back (para_finish.f90:222-228).
program portal_f90_demo
  use iso_c_binding
  use portal_hip
  implicit none
  character(len=512) :: fin, fout, fdat
  integer :: node
  integer(c_int) :: node_num, bulk_node_num, e_num, r_num, c_num, c_temp, c_par, kmax, maxne, n_sphere, nvel, rc
  real(c_double) :: tx, dx
  real(c_double), allocatable, target :: XYZ(:,:), vel_all(:,:), temp_all(:,:), b_all(:), h2(:), ab(:), dust(:)
  real(c_double), allocatable, target :: lev_p(:,:), e_levs(:), t_Aji(:), nu(:), t_Bij(:), kd_v(:), t_Cji(:,:,:)
  real(c_double), allocatable, target :: col_temps(:), xth(:,:), gvec(:), bvec_all(:,:), pol_pops(:,:)
  real(c_double), allocatable, target :: pol_rates(:,:,:), JJ_20(:,:)
  integer(c_int), allocatable :: del_ne(:,:), del_num_ne(:), j_lev(:), r_trans(:,:), c_trans(:,:)
  type(c_ptr) :: ctx
  type(portal_se_params) :: p
  type(portal_stats) :: st

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, fdat)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) node_num, bulk_node_num, e_num, r_num, c_num, c_temp, c_par, kmax, maxne, n_sphere, nvel
  read(10) tx, dx
  allocate(XYZ(3,node_num), vel_all(3,node_num), temp_all(2,node_num), b_all(node_num), h2(node_num))
  allocate(ab(node_num), dust(node_num), lev_p(e_num,node_num), del_ne(node_num,maxne), del_num_ne(node_num))
  allocate(j_lev(e_num), e_levs(e_num), r_trans(2,r_num), t_Aji(r_num), nu(r_num), t_Bij(r_num), kd_v(r_num))
  allocate(c_trans(2,c_num), t_Cji(c_temp,c_num,c_par), col_temps(c_temp), xth(3,n_sphere), gvec(3))
  allocate(bvec_all(3,node_num), pol_pops(node_num,2*e_num), pol_rates(node_num,3,2*e_num), JJ_20(node_num,r_num))
  read(10) XYZ, vel_all, temp_all, b_all, h2, ab, dust, lev_p, del_ne, del_num_ne
  read(10) j_lev, e_levs, r_trans, t_Aji, nu, t_Bij, kd_v, c_trans, t_Cji, col_temps
  read(10) xth, gvec, bvec_all
  close(10)

  rc = portal_ctx_create(ctx, 0)
  if (rc /= PORTAL_OK) stop 'portal_ctx_create failed'
  rc = portal_upload_grid(ctx, node_num, bulk_node_num, XYZ, vel_all, temp_all, b_all, h2, ab, dust, &
                          del_ne, del_num_ne, maxne, tx)
  if (rc == PORTAL_OK) rc = portal_upload_species(ctx, e_num, j_lev, e_levs, r_num, r_trans, t_Aji, nu, t_Bij, &
                          kd_v, c_num, c_trans, c_temp, c_par, t_Cji, col_temps, kmax)
  if (rc == PORTAL_OK) rc = portal_set_levpop(ctx, lev_p)
  p%n_sphere = n_sphere
  p%xth = c_loc(xth)
  p%nvel = nvel
  p%dx = dx
  p%gvec = c_loc(gvec)
  p%bvec_all = c_loc(bvec_all)
  p%n_depth = 4096
  p%extra_info = 0
  p%precision = PORTAL_FP64
  p%star_mod = 0
  p%nth = 0
  p%nph = 0
  p%R_star = 0.d0
  p%T_star = 0.d0
  p%X_star = c_null_ptr
  pol_pops = 0.d0
  if (rc == PORTAL_OK) rc = portal_se_pass(ctx, p, pol_pops, pol_rates, JJ_20, st)
  if (rc /= PORTAL_OK) then
    write(*,*) 'portal error ', rc, trim(portal_last_error(ctx))
    stop 1
  endif
  write(*,'(a,f10.3,a,i12)') 'se_pass device ms:', st%t_total_ms, '  crossings:', st%n_cross
  rc = portal_ctx_destroy(ctx)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) pol_pops
  close(11)
  if (len_trim(fdat) > 0) then
    open(unit=25, file=trim(fdat), recl=8824)
    do node = 1, node_num
      write(25,*) pol_pops(node,:)
    end do
    close(25)
  endif
end program
