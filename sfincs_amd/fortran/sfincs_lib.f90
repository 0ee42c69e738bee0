! Fortran host of the MI355X-native SFINCS step.
!
! Keeps the reference's model-driver entry points (sfincs_lib.f90:85-819:
! sfincs_initialize / sfincs_update(dtrange) / sfincs_finalize, public t and dt) and the
! BMI C names (sfincs_bmi.f90:56-387), and calls the HIP kernels through the C ABI of
! include/sfx.h with iso_c_binding. The host reads the reference's input deck:
!   sfincs.inp (key = value, reference defaults), inputformat bin (indexfile/mskfile/
!   depfile, stream binary) or asc (mskfile/depfile, nmax rows of mmax values), or a
!   legacy binary qtrfile (sfincs_quadtree.F90:210-292) with a binary mskfile, or a NetCDF
!   qtrfile (:295-442); sbgfile as NetCDF or legacy binary (sfincs_subgrid.F90:20-53);
!   bndfile/bzsfile, srcfile/disfile, prcfile, wndfile (ASCII time series).
module sfx_capi
   use iso_c_binding
   implicit none
   type, bind(C) :: sfx_params
      real(c_float) :: t0, t1, tspinup, alpha, theta, hmin_cfl, manning, manning_land, manning_sea
      real(c_float) :: rgh_lev_land, zsini, dtmax, huthresh, huvmin, rhoa, rhow, advlim, slopelim
      real(c_float) :: btfilter, btrelax, wiggle_factor, wiggle_threshold, uvlim, uvmax, nuvisc
      real(c_float) :: nuviscfac, latitude, g, pi
      integer(c_int32_t) :: advection, advection_scheme, viscosity, coriolis, friction2d
      integer(c_int32_t) :: advection_mask, wiggle_suppression, store_zsmax, bndtype, h73table
      integer(c_int32_t) :: store_velmax, store_fluxmax, store_twet
      real(c_float) :: twet_threshold
      integer(c_int32_t) :: timestep_analysis, use_qext
      real(c_float) :: structure_relax, qinf, qinf_zmin
      real(c_float) :: dtwnd
      integer(c_int32_t) :: baro
      real(c_float) :: pavbnd, gapres
      integer(c_int32_t) :: spinup_meteo, amprblock, usespwprecip
      real(c_float) :: spw_merge_frac, factor_wind, factor_pres, factor_prcp, factor_spw_size
      real(c_float) :: sfacinf
   end type
   type, bind(C) :: sfx_quadtree_input
      integer(c_int32_t) :: nr_points, nr_levels
      real(c_float) :: x0, y0, dx, dy, rotation_deg
      type(c_ptr) :: level, md, mu, nd, nu, mask
      type(c_ptr) :: n, m, md1, md2, mu1, mu2, nd1, nd2, nu1, nu2
   end type
   type, bind(C) :: sfx_step_stats
      real(c_double) :: t
      real(c_float) :: dt, min_dt
      integer(c_int64_t) :: nt
      integer(c_int32_t) :: error
      real(c_double) :: cell_updates, seconds
   end type
   interface
      subroutine sfx_params_default(p) bind(C, name='sfx_params_default')
         import :: sfx_params
         type(sfx_params), intent(out) :: p
      end subroutine
      integer(c_int) function sfx_model_create_regular(p, nmax, mmax, x0, y0, dx, dy, rot, msk, out) &
            bind(C, name='sfx_model_create_regular')
         import :: sfx_params, c_int, c_int32_t, c_float, c_ptr
         type(sfx_params), intent(in) :: p
         integer(c_int32_t), value :: nmax, mmax
         real(c_float), value :: x0, y0, dx, dy, rot
         type(c_ptr), value :: msk
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function sfx_model_create_quadtree(p, qt, out) bind(C, name='sfx_model_create_quadtree')
         import :: sfx_params, sfx_quadtree_input, c_int, c_ptr
         type(sfx_params), intent(in) :: p
         type(sfx_quadtree_input), intent(in) :: qt
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function sfx_model_table_bytes(m, name, nbytes) bind(C, name='sfx_model_table_bytes')
         import :: c_int, c_ptr, c_char, c_size_t
         type(c_ptr), value :: m
         character(kind=c_char), intent(in) :: name(*)
         integer(c_size_t), intent(out) :: nbytes
      end function
      integer(c_int) function sfx_model_get_table(m, name, dst, nbytes) bind(C, name='sfx_model_get_table')
         import :: c_int, c_ptr, c_char, c_size_t
         type(c_ptr), value :: m
         character(kind=c_char), intent(in) :: name(*)
         type(c_ptr), value :: dst
         integer(c_size_t), value :: nbytes
      end function
      integer(c_int) function sfx_model_set_bathymetry(m, zb) bind(C, name='sfx_model_set_bathymetry')
         import :: c_int, c_ptr
         type(c_ptr), value :: m, zb
      end function
      integer(c_int) function sfx_model_set_boundary(m, nbnd, x, y, nt, t, zs) bind(C, name='sfx_model_set_boundary')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: nbnd, nt
         type(c_ptr), value :: x, y, t, zs
      end function
      integer(c_int) function sfx_model_set_sources(m, nsrc, x, y, nt, t, q) bind(C, name='sfx_model_set_sources')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: nsrc, nt
         type(c_ptr), value :: x, y, t, q
      end function
      integer(c_int) function sfx_model_set_boundary_ig(m, zsi) bind(C, name='sfx_model_set_boundary_ig')
         import :: c_int, c_ptr
         type(c_ptr), value :: m, zsi
      end function
      integer(c_int) function sfx_model_set_river_boundary(m, nbdr, x, y, xi, yi, slope, dist) &
            bind(C, name='sfx_model_set_river_boundary')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: nbdr
         type(c_ptr), value :: x, y, xi, yi, slope, dist
      end function
      integer(c_int) function sfx_model_set_drainage(m, ndrn, xi, yi, xo, yo, ty, par) &
            bind(C, name='sfx_model_set_drainage')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: ndrn
         type(c_ptr), value :: xi, yi, xo, yo, ty, par
      end function
      integer(c_int) function sfx_model_set_thin_dams(m, npoly, nv, x, y) bind(C, name='sfx_model_set_thin_dams')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: npoly
         type(c_ptr), value :: nv, x, y
      end function
      integer(c_int) function sfx_model_set_infiltration(m, q) bind(C, name='sfx_model_set_infiltration')
         import :: c_int, c_ptr
         type(c_ptr), value :: m, q
      end function
      integer(c_int) function sfx_model_set_curve_number(m, s) bind(C, name='sfx_model_set_curve_number')
         import :: c_int, c_ptr
         type(c_ptr), value :: m, s
      end function
      integer(c_int) function sfx_model_set_weirs(m, npoly, nv, x, y, par) bind(C, name='sfx_model_set_weirs')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: npoly
         type(c_ptr), value :: nv, x, y, par
      end function
      integer(c_int) function sfx_model_set_rain(m, nt, t, v) bind(C, name='sfx_model_set_rain')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: nt
         type(c_ptr), value :: t, v
      end function
      integer(c_int) function sfx_model_set_wind(m, nt, t, mag, dir) bind(C, name='sfx_model_set_wind')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: nt
         type(c_ptr), value :: t, mag, dir
      end function
      integer(c_int) function sfx_model_find_cells(m, n, x, y, nm) bind(C, name='sfx_model_find_cells')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: n
         type(c_ptr), value :: x, y, nm
      end function
      integer(c_int) function sfx_model_read_meteo_file(m, kind, path, tref) bind(C, name='sfx_model_read_meteo_file')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: m
         character(kind=c_char), intent(in) :: kind(*), path(*), tref(*)
      end function
      integer(c_int) function sfx_model_read_subgrid_binary(m, path) bind(C, name='sfx_model_read_subgrid_binary')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: m
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function sfx_model_read_subgrid_netcdf(m, path) bind(C, name='sfx_model_read_subgrid_netcdf')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: m
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function sfx_model_create_quadtree_netcdf(p, path, with_bed, out) &
            bind(C, name='sfx_model_create_quadtree_netcdf')
         import :: sfx_params, c_int, c_int32_t, c_ptr, c_char
         type(sfx_params), intent(in) :: p
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int32_t), value :: with_bed
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function sfx_update_zbuvmx(h) bind(C, name='sfx_update_zbuvmx')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function sfx_model_initialize(m) bind(C, name='sfx_model_initialize')
         import :: c_int, c_ptr
         type(c_ptr), value :: m
      end function
      integer(c_int) function sfx_model_set_restart(m, path) bind(C, name='sfx_model_set_restart')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: m
         character(kind=c_char), intent(in) :: path(*)
      end function
      subroutine sfx_model_destroy(m) bind(C, name='sfx_model_destroy')
         import :: c_ptr
         type(c_ptr), value :: m
      end subroutine
      integer(c_int) function sfx_create(m, device, out) bind(C, name='sfx_create')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: m
         integer(c_int32_t), value :: device
         type(c_ptr), intent(out) :: out
      end function
      integer(c_int) function sfx_destroy(h) bind(C, name='sfx_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function
      integer(c_int) function sfx_update(h, dtrange, st) bind(C, name='sfx_update')
         import :: c_int, c_ptr, c_double, sfx_step_stats
         type(c_ptr), value :: h
         real(c_double), value :: dtrange
         type(sfx_step_stats), intent(out) :: st
      end function
      integer(c_int) function sfx_update_to(h, tend, st) bind(C, name='sfx_update_to')
         import :: c_int, c_ptr, c_double, sfx_step_stats
         type(c_ptr), value :: h
         real(c_double), value :: tend
         type(sfx_step_stats), intent(out) :: st
      end function
      integer(c_int) function sfx_write_restart(h, path) bind(C, name='sfx_write_restart')
         import :: c_int, c_ptr, c_char
         type(c_ptr), value :: h
         character(kind=c_char), intent(in) :: path(*)
      end function
      integer(c_int) function sfx_get_state(h, name, dst, nbytes) bind(C, name='sfx_get_state')
         import :: c_int, c_ptr, c_char, c_size_t
         type(c_ptr), value :: h
         character(kind=c_char), intent(in) :: name(*)
         type(c_ptr), value :: dst
         integer(c_size_t), value :: nbytes
      end function
      integer(c_int) function sfx_set_state(h, name, src, nbytes) bind(C, name='sfx_set_state')
         import :: c_int, c_ptr, c_char, c_size_t
         type(c_ptr), value :: h
         character(kind=c_char), intent(in) :: name(*)
         type(c_ptr), value :: src
         integer(c_size_t), value :: nbytes
      end function
      integer(c_int) function sfx_model_timestep_analysis_per_cell(m, tsum, twet, tlim, nt, dtc, pct) &
            bind(C, name='sfx_model_timestep_analysis_per_cell')
         import :: c_int, c_ptr, c_int64_t
         type(c_ptr), value :: m, tsum, twet, tlim, dtc, pct
         integer(c_int64_t), value :: nt
      end function
      integer(c_int) function sfx_model_find_cell(m, x, y, idx) bind(C, name='sfx_model_find_cell')
         import :: c_int, c_ptr, c_float, c_int32_t
         type(c_ptr), value :: m
         real(c_float), value :: x, y
         integer(c_int32_t), intent(out) :: idx
      end function
      integer(c_int) function sfx_get_time(h, t, dt, nt) bind(C, name='sfx_get_time')
         import :: c_int, c_ptr, c_double, c_float, c_int64_t
         type(c_ptr), value :: h
         real(c_double), intent(out) :: t
         real(c_float), intent(out) :: dt
         integer(c_int64_t), intent(out) :: nt
      end function
      integer(c_int) function sfx_last_error(buf, n) bind(C, name='sfx_last_error')
         import :: c_int, c_char, c_size_t
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_size_t), value :: n
      end function
   end interface
end module sfx_capi


module sfincs_lib
   use iso_c_binding
   use sfx_capi
   implicit none
   private
   public :: sfincs_initialize, sfincs_update, sfincs_finalize, t, dt, sfincs_read_model, host_model
   public :: host_np, host_npuv, host_ncuv, host_zs, host_zsmax, host_zb, host_sync, host_t0, host_t1
   public :: host_xz, host_yz, host_zmin, host_area, host_iref, host_mesh_fields, host_last_error, host_find_cell
   public :: host_qext_requested, host_qext_on, host_qext, host_qext_switch
   public :: host_expose, host_update_zbuv, host_zn, host_zm

   real*8, target :: t = 0.0d0
   real*4, target :: dt = 0.0
   real*8 :: host_t0 = 0.0d0, host_t1 = 0.0d0
   integer :: host_np = 0, host_npuv = 0, host_ncuv = 0
   integer :: host_flags(4) = 0   ! store_velmax, store_fluxmax, store_twet, timestep_analysis
   ! external sources (qext, m/s per cell): requested by the BMI initialize as
   ! sfincs_bmi.f90:60 does, written by a coupler through get_value_ptr, uploaded before
   ! each update while switched on (set_logical)
   logical :: host_qext_requested = .false., host_qext_on = .false.
   real*4, allocatable, target :: host_qext(:)
   real*8, allocatable, target :: host_zs(:)
   real*4, allocatable, target :: host_zsmax(:), host_zb(:)
   ! get_value_ptr hands out host copies of zs and zb, which a coupler may write, as it
   ! writes the reference's own arrays (sfincs_bmi.f90:95-129): once handed out, a copy is
   ! compared with its shadow before each update and uploaded when it changed, and zs is
   ! copied back after each update so that the pointer keeps showing the model's levels
   logical :: host_zs_out = .false., host_zb_out = .false.
   real*8, allocatable :: shadow_zs(:)
   real*4, allocatable :: shadow_zb(:)
   integer(c_int32_t), allocatable :: host_zn(:), host_zm(:)
   ! mesh fields for the BMI (cell centres, bed or lowest subgrid level, cell areas)
   real*4, allocatable, target :: host_xz(:), host_yz(:), host_zmin(:), host_area(:)
   integer(c_int8_t), allocatable, target :: host_iref(:)
   type(c_ptr) :: model = c_null_ptr
   type(c_ptr) :: handle = c_null_ptr
   integer :: logu = 0
   ! scheduled output (sfincs_output.f90:8-60, 80-258; triggers sfincs_lib.f90:419-489):
   ! binary map (zs.dat, z_volume.dat), maximum (zsmax.dat, velmax.dat, tmax.dat, qmax.dat)
   ! and restart files; NetCDF output would need a library this host does not link
   real*8 :: tmapout = 1.0d9, tmaxout = 1.0d9, trstout = 1.0d9
   real*8 :: out_dtmap = 0.0d0, out_dtmax = 0.0d0, out_dtrst = 0.0d0, out_trst = -999.0d0, out_t0 = 0.0d0
   real*8 :: host_tref_sec = 0.0d0
   integer :: ntmapout = 0, ntmaxout = 0
   integer :: u_zs = 0, u_zv = 0, u_zsmax = 0, u_vmax = 0, u_twet = 0, u_qmax = 0
   logical :: out_zvolume = .false., host_subgrid = .false.
   integer :: host_store_zsmax = 0
   real*4 :: host_huthresh = 0.0
   ! history output at observation points (read_obs_points, sfincs_obspoints.f90:5-135;
   ! open_his_output / write_his_output, sfincs_output.f90:567-655): zst.txt, prcp.txt
   ! (with a rain series) and qdrain.txt (drainage structures) every dthisout from t0
   real*8 :: thisout = 1.0d9, out_dthis = 600.0d0
   integer :: nthisout = 0, nobs = 0, his_nsrc = 0, his_ndrn = 0
   integer(c_int32_t), allocatable, target :: obs_nm(:)
   real(c_float), allocatable, target :: obs_x(:), obs_y(:)
   real*4, allocatable :: his_tp(:), his_vp(:)
   logical :: his_qdrain = .false.

contains

   subroutine logmsg(s)
      character(*), intent(in) :: s
      write(*, '(a)') trim(s)
      if (logu /= 0) write(logu, '(a)') trim(s)
   end subroutine

   subroutine fail(where)
      character(*), intent(in) :: where
      character(kind=c_char) :: buf(512)
      character(512) :: msg
      integer :: i, rc
      rc = sfx_last_error(buf, int(512, c_size_t))
      msg = ' '
      do i = 1, 512
         if (buf(i) == c_null_char) exit
         msg(i:i) = buf(i)
      enddo
      call logmsg('Error   : '//trim(where)//': '//trim(msg))
   end subroutine

   ! ---- sfincs.inp ------------------------------------------------------------------
   subroutine inp_get(key, val, found)
      character(*), intent(in) :: key
      character(256), intent(out) :: val
      logical, intent(out) :: found
      character(512) :: line
      character(128) :: k
      integer :: u, ios, ieq
      found = .false.
      val = ' '
      open(newunit=u, file='sfincs.inp', status='old', action='read', iostat=ios)
      if (ios /= 0) return
      do
         read(u, '(a)', iostat=ios) line
         if (ios /= 0) exit
         ieq = index(line, '=')
         if (ieq == 0) cycle
         k = adjustl(line(1:ieq - 1))
         if (trim(k) == key) then
            val = adjustl(line(ieq + 1:))
            found = .true.
            exit
         endif
      enddo
      close(u)
   end subroutine

   subroutine inp_real(key, v)
      character(*), intent(in) :: key
      real(c_float), intent(inout) :: v
      character(256) :: s
      logical :: f
      call inp_get(key, s, f)
      if (f) read(s, *) v
   end subroutine

   subroutine inp_int(key, v)
      character(*), intent(in) :: key
      integer, intent(inout) :: v
      character(256) :: s
      logical :: f
      call inp_get(key, s, f)
      if (f) read(s, *) v
   end subroutine

   subroutine inp_str(key, v, default)
      character(*), intent(in) :: key, default
      character(256), intent(out) :: v
      logical :: f
      call inp_get(key, v, f)
      if (.not. f) v = default
   end subroutine

   subroutine inp_flag(key, v)
      character(*), intent(in) :: key
      integer(c_int32_t), intent(inout) :: v
      character(256) :: s
      logical :: f
      call inp_get(key, s, f)
      if (f) then
         select case (s(1:1))
         case ('t', 'T', '1', 'y', 'Y')
            v = 1
         case default
            v = 0
         end select
      endif
   end subroutine

   ! seconds since 0001-01-01 for 'yyyymmdd hhmmss' (proleptic Gregorian)
   real*8 function date_seconds(s)
      character(*), intent(in) :: s
      integer :: y, mo, d, h, mi, se, a, yy, mm, jdn
      read(s(1:4), *) y; read(s(5:6), *) mo; read(s(7:8), *) d
      read(s(10:11), *) h; read(s(12:13), *) mi; read(s(14:15), *) se
      a = (14 - mo) / 12
      yy = y + 4800 - a
      mm = mo + 12 * a - 3
      jdn = d + (153 * mm + 2) / 5 + 365 * yy + yy / 4 - yy / 100 + yy / 400 - 32045
      date_seconds = dble(jdn) * 86400.0d0 + dble(h) * 3600.0d0 + dble(mi) * 60.0d0 + dble(se)
   end function

   ! NetCDF files start with 'CDF' (classic) or the HDF5 signature (NetCDF-4).
   logical function is_netcdf(fname)
      character(*), intent(in) :: fname
      character(4) :: magic
      integer :: u, ios
      is_netcdf = .false.
      magic = '    '
      open(newunit=u, file=trim(fname), form='unformatted', access='stream', status='old', action='read', iostat=ios)
      if (ios /= 0) return
      read(u, iostat=ios) magic
      close(u)
      if (ios /= 0) return
      is_netcdf = magic(1:3) == 'CDF' .or. magic(2:4) == 'HDF'
   end function

   integer function count_lines(fname)
      character(*), intent(in) :: fname
      character(1024) :: line
      integer :: u, ios
      count_lines = 0
      open(newunit=u, file=trim(fname), status='old', action='read', iostat=ios)
      if (ios /= 0) then
         count_lines = -1
         return
      endif
      do
         read(u, '(a)', iostat=ios) line
         if (ios /= 0) exit
         if (len_trim(line) > 0) count_lines = count_lines + 1
      enddo
      close(u)
   end function

   ! Polylines in the tekal block format of thd and weir files (sfincs_structures.f90:
   ! 113-160, 421-470): per polyline a name line, a line "nrows ncols", then nrows lines of
   ! x y and npar parameters. Returns the vertex count per polyline, the coordinates and the
   ! parameters (npar per vertex, vertex-major), or ok = .false. on a malformed file.
   subroutine read_polylines(fname, npar, nv, xv, yv, pv, ok)
      character(*), intent(in) :: fname
      integer, intent(in) :: npar
      integer(c_int32_t), allocatable, intent(out) :: nv(:)
      real(c_float), allocatable, intent(out) :: xv(:), yv(:), pv(:)
      logical, intent(out) :: ok
      character(256) :: name
      integer :: u, ios, npoly, nvert, nrows, ncols, i, k, j
      real(c_float) :: row(2 + 8)
      ok = .false.
      open(newunit=u, file=trim(fname), status='old', action='read', iostat=ios)
      if (ios /= 0) return
      npoly = 0
      nvert = 0
      do
         read(u, *, iostat=ios) name
         if (ios /= 0) exit
         read(u, *, iostat=ios) nrows, ncols
         if (ios /= 0) exit
         if (nrows < 2 .or. ncols < 2 + npar) then
            close(u)
            return
         endif
         do i = 1, nrows
            read(u, *, iostat=ios) row(1)
            if (ios /= 0) then
               close(u)
               return
            endif
         enddo
         npoly = npoly + 1
         nvert = nvert + nrows
      enddo
      rewind(u)
      allocate(nv(npoly), xv(nvert), yv(nvert), pv(max(1, npar * nvert)))
      pv = 0.0
      k = 0
      do i = 1, npoly
         read(u, *) name
         read(u, *) nrows, ncols
         nv(i) = nrows
         do j = 1, nrows
            k = k + 1
            read(u, *) row(1:2 + npar)
            xv(k) = row(1)
            yv(k) = row(2)
            if (npar > 0) pv(npar * (k - 1) + 1:npar * k) = row(3:2 + npar)
         enddo
      enddo
      close(u)
      ok = .true.
   end subroutine

   ! ---- sfincs_initialize (sfincs_lib.f90:85-335) --------------------------------------
   ! Reads the input deck into an initialized host model, then moves it to HBM.
   function sfincs_initialize() result(ierr)
      integer :: ierr, i
      integer(c_int) :: rc
      ierr = sfincs_read_model()
      if (ierr /= 0) return
      call host_mesh_fields()
      call table_size('zs', 8, host_np)
      if (host_qext_requested) then
         if (allocated(host_qext)) deallocate(host_qext)
         allocate(host_qext(host_np))
         host_qext = 0.0
         host_qext_on = .true.
      endif
      call table_size('uv_index_z_nm', 4, host_npuv)
      if (allocated(host_zs)) deallocate(host_zs, host_zsmax)
      allocate(host_zs(host_np), host_zsmax(host_np))
      host_zs_out = .false.
      host_zb_out = .false.
      ! the replacement of initialize_openacc: tables and state to HBM once
      rc = sfx_create(model, 0_c_int32_t, handle)
      if (rc < 0) then
         call fail('sfx_create')
         ierr = -1
         return
      endif
      t = host_t0
      dt = 1.0e-6
      call table_size('nmindsrc', 4, i)
      call table_size('drainage_type', 4, his_ndrn)
      his_nsrc = i - 2 * his_ndrn
      call output_init()
      call logmsg('---------- Starting simulation -----------')
      ierr = 0
   end function

   ! Input deck (sfincs.inp and the files it names) -> initialized host model; no GPU.
   function sfincs_read_model() result(ierr)
      integer :: ierr
      type(sfx_params), target :: p
      integer :: nmax, mmax, np, ios, u, i, j, n, m, nt, nb, rc
      real(c_float) :: x0, y0, dx, dy, rot, dtmaxout, tdum
      character(256) :: tref, tstart, tstop, inputformat, depfile, mskfile, indexfile, qtrfile
      character(256) :: bndfile, bzsfile, bzifile, srcfile, disfile, prcfile, wndfile, rstfile, bdrfile, s
      character(256) :: sbgfile, drnfile, thdfile, weirfile, qinffile, obsfile, scsfile
      character(256) :: metfile(5)
      character(4), parameter :: metkind(5) = ['amu ', 'amv ', 'amp ', 'ampr', 'spw ']
      character(8), parameter :: metkey(5) = ['amufile ', 'amvfile ', 'ampfile ', 'amprfile', 'spwfile ']
      real(c_float) :: rdum
      integer(c_int32_t), allocatable, target :: pl_nv(:)
      real(c_float), allocatable, target :: pl_x(:), pl_y(:), pl_p(:)
      logical :: pl_ok
      character(512) :: line
      integer(c_int32_t), allocatable, target :: drn_type(:)
      real(c_float), allocatable, target :: drn_par(:,:)
      integer(c_int8_t), allocatable, target :: msk(:), mskg(:,:)
      integer(c_int32_t), allocatable :: indices(:)
      real(c_float), allocatable, target :: zbg(:,:), zb(:), xb(:), yb(:), tb(:), zsb(:,:), tmp(:), zsi(:,:)
      real(c_float), allocatable, target :: tt(:), v1(:), v2(:)
      integer(c_size_t) :: nbytes
      logical :: f
      ierr = 0
      ! a second initialize in the same process starts from a clean host
      if (logu /= 0) close(logu)
      logu = 0
      if (c_associated(model)) call sfx_model_destroy(model)
      model = c_null_ptr
      open(newunit=logu, file='sfincs_mi355x.log', status='replace')
      call logmsg('------------ SFINCS (MI355X-native time step) ------------')
      call sfx_params_default(p)
      call inp_str('tref', tref, '20000101 000000')
      call inp_str('tstart', tstart, '20000101 000000')
      call inp_str('tstop', tstop, '20000101 000000')
      call inp_get('tref', s, f)
      if (.not. f) tref = tstart
      p%t0 = real(date_seconds(tstart) - date_seconds(tref), c_float)
      p%t1 = real(date_seconds(tstop) - date_seconds(tref), c_float)
      call inp_real('tspinup', p%tspinup)
      call inp_real('alpha', p%alpha); call inp_real('theta', p%theta); call inp_real('hmin_cfl', p%hmin_cfl)
      call inp_real('manning', p%manning); call inp_real('manning_land', p%manning_land)
      call inp_real('manning_sea', p%manning_sea); call inp_real('rgh_lev_land', p%rgh_lev_land)
      call inp_real('zsini', p%zsini); call inp_real('dtmax', p%dtmax); call inp_real('huthresh', p%huthresh)
      call inp_real('huvmin', p%huvmin); call inp_real('rhoa', p%rhoa); call inp_real('rhow', p%rhow)
      call inp_real('advlim', p%advlim); call inp_real('slopelim', p%slopelim); call inp_real('btfilter', p%btfilter)
      call inp_real('btrelax', p%btrelax); call inp_real('wiggle_factor', p%wiggle_factor)
      call inp_real('wiggle_threshold', p%wiggle_threshold); call inp_real('uvlim', p%uvlim)
      call inp_real('uvmax', p%uvmax); call inp_real('nuvisc', p%nuvisc); call inp_real('nuviscfac', p%nuviscfac)
      call inp_real('latitude', p%latitude)
      call inp_flag('viscosity', p%viscosity); call inp_flag('coriolis', p%coriolis)
      call inp_flag('friction2d', p%friction2d); call inp_flag('advection_mask', p%advection_mask)
      call inp_flag('wiggle_suppression', p%wiggle_suppression); call inp_flag('h73table', p%h73table)
      i = p%advection
      call inp_int('advection', i); p%advection = merge(1, 0, i > 0)
      call inp_str('advection_scheme', s, 'upw1')
      if (trim(s) == 'original') p%advection_scheme = 0
      ! maximum and store switches (sfincs_input.f90:64, 277-285, 449-463, 501-504)
      dtmaxout = 9999999.0
      call inp_real('dtmaxout', dtmaxout)
      p%store_zsmax = merge(1, 0, dtmaxout > 0.0)
      i = 0
      call inp_int('storevelmax', i); p%store_velmax = merge(1, 0, i == 1 .and. dtmaxout > 0.0)
      i = 0
      call inp_int('storefluxmax', i); p%store_fluxmax = merge(1, 0, i == 1 .and. dtmaxout > 0.0)
      i = 0
      call inp_int('storetwet', i); p%store_twet = merge(1, 0, i == 1)
      call inp_real('twet_threshold', p%twet_threshold)
      call inp_flag('timestep_analysis', p%timestep_analysis)
      ! output times (sfincs_input.f90:61-66, 341-343, 441-444, 577-582)
      out_dtmap = 0.0d0
      tdum = 0.0
      call inp_real('dtout', tdum)
      if (tdum == 0.0) call inp_real('dtmapout', tdum)
      out_dtmap = tdum
      out_dtmax = dtmaxout
      tdum = 0.0
      call inp_real('dtrstout', tdum)
      out_dtrst = tdum
      tdum = -999.0
      call inp_real('trstout', tdum)
      out_trst = tdum
      tdum = -999.0
      call inp_real('t0out', tdum)
      out_t0 = tdum
      if (out_t0 < -900.0d0) out_t0 = p%t0
      out_t0 = max(out_t0, dble(p%t0))
      i = 0
      call inp_int('storezvolume', i)
      out_zvolume = i == 1
      host_tref_sec = date_seconds(tref)
      host_store_zsmax = p%store_zsmax
      call inp_str('outputformat', s, 'net')
      if (s(1:3) == 'net' .and. (out_dtmap > 1.0d-6 .or. out_dtmax > 1.0d-6)) &
         call logmsg('Info    : NetCDF output is not available in this host; map and maximum output are binary')
      host_flags = [p%store_velmax, p%store_fluxmax, p%store_twet, p%timestep_analysis]
      if (host_qext_requested) p%use_qext = 1
      host_t0 = p%t0; host_t1 = p%t1
      x0 = 0.0; y0 = 0.0; dx = 0.0; dy = 0.0; rot = 0.0; nmax = 0; mmax = 0
      call inp_int('nmax', nmax); call inp_int('mmax', mmax)
      call inp_real('x0', x0); call inp_real('y0', y0); call inp_real('dx', dx); call inp_real('dy', dy)
      call inp_real('rotation', rot)
      call inp_str('inputformat', inputformat, 'bin')
      call inp_str('depfile', depfile, 'none'); call inp_str('mskfile', mskfile, 'none')
      call inp_str('indexfile', indexfile, 'none'); call inp_str('qtrfile', qtrfile, 'none')
      call inp_str('bndfile', bndfile, 'none'); call inp_str('bzsfile', bzsfile, 'none')
      call inp_str('srcfile', srcfile, 'none'); call inp_str('disfile', disfile, 'none')
      call inp_str('prcfile', prcfile, 'none'); call inp_str('wndfile', wndfile, 'none')
      call inp_str('rstfile', rstfile, 'none'); call inp_str('bdrfile', bdrfile, 'none')
      call inp_str('bzifile', bzifile, 'none')
      call inp_str('sbgfile', sbgfile, 'none')
      call inp_str('drnfile', drnfile, 'none')
      call inp_str('thdfile', thdfile, 'none'); call inp_str('weirfile', weirfile, 'none')
      call inp_real('structure_relax', p%structure_relax)
      ! constant infiltration (sfincs_input.f90:78,100,248): qinf in mm/h, converted by the model
      call inp_real('qinf', p%qinf); call inp_real('qinf_zmin', p%qinf_zmin)
      call inp_str('qinffile', qinffile, 'none')
      ! curve-number infiltration, method A (sfincs_input.f90:102; sfincs_infiltration.f90:106-112)
      call inp_str('scsfile', scsfile, 'none'); call inp_real('sfacinf', p%sfacinf)
      call inp_str('infiltrationfile', s, 'none')
      if (s(1:4) /= 'none') then
         call logmsg('Error   : '//trim(s)//' is a NetCDF infiltration file; this host reads qinf and qinffile only')
         ierr = -1
         return
      endif
      ! history output (sfincs_input.f90:67, 274, 289)
      tdum = 600.0
      call inp_real('dthisout', tdum)
      out_dthis = tdum
      call inp_str('obsfile', obsfile, 'none')
      i = 1; call inp_int('storeqdrain', i); his_qdrain = i /= 0
      ! gridded and spiderweb meteo (sfincs_input.f90:69,92-113,200-203,236-247): the ASCII files
      call inp_real('dtwnd', p%dtwnd); i = p%baro; call inp_int('baro', i); p%baro = i
      call inp_real('pavbnd', p%pavbnd); call inp_real('gapres', p%gapres)
      i = 0; call inp_int('spinup_meteo', i); p%spinup_meteo = merge(1, 0, i /= 0)
      i = 1; call inp_int('amprblock', i); p%amprblock = merge(1, 0, i /= 0)
      i = 1; call inp_int('usespwprecip', i); p%usespwprecip = merge(1, 0, i /= 0)
      call inp_real('spwmergefrac', p%spw_merge_frac)
      call inp_real('factor_wind', p%factor_wind); call inp_real('factor_pres', p%factor_pres)
      call inp_real('factor_prcp', p%factor_prcp); call inp_real('factor_spw_size', p%factor_spw_size)
      do j = 1, 5
         call inp_str(trim(metkey(j)), metfile(j), 'none')
      enddo
      call inp_str('netamuamvfile', s, 'none')
      if (s(1:4) == 'none') call inp_str('netampfile', s, 'none')
      if (s(1:4) == 'none') call inp_str('netamprfile', s, 'none')
      if (s(1:4) == 'none') call inp_str('netspwfile', s, 'none')
      if (s(1:4) /= 'none') then
         call logmsg('Error   : '//trim(s)//' is NetCDF meteo input; this host reads the ASCII amu/amv/amp/ampr '// &
                     'and spw files')
         ierr = -1
         return
      endif
      call inp_str('z0lfile', s, 'none')
      i = 0; call inp_int('crsgeo', i)
      rdum = -999.0; call inp_real('waveage', rdum)
      call inp_str('utmzone', line, 'nil')
      if (s(1:4) /= 'none' .or. i /= 0 .or. rdum > 0.0 .or. line(1:3) /= 'nil') then
         call logmsg('Error   : z0lfile, crsgeo, waveage and utmzone are not supported by this host')
         ierr = -1
         return
      endif
      ! a map type on a quadtree mesh (any type but 'con', which a qinf > 0 selects first;
      ! sfincs_infiltration.f90:92-133,172-186)
      if ((qinffile(1:4) /= 'none' .or. scsfile(1:4) /= 'none') .and. qtrfile(1:4) /= 'none' .and. &
          p%qinf <= 0.0) then
         call logmsg('Error ! Infiltration input for quadtree mesh model can only be specified using the '// &
                     'infiltrationfile Netcdf format! !')
         ierr = -1
         return
      endif

      if (qtrfile(1:4) /= 'none' .and. index(qtrfile, '.nc') > 0) then
         ! NetCDF quadtree file, chosen by its name as quadtree_read_file does
         ! (sfincs_quadtree.F90:85-95); the mask is one of its variables
         call logmsg('Info    : reading QuadTree netCDF file '//trim(qtrfile))
         rc = sfx_model_create_quadtree_netcdf(p, trim(qtrfile)//c_null_char, &
                                               merge(1_c_int32_t, 0_c_int32_t, sbgfile(1:4) == 'none'), model)
         if (rc < 0) then
            call fail('quadtree file')
            ierr = -1
            return
         endif
      elseif (qtrfile(1:4) /= 'none') then
         call read_quadtree(qtrfile, mskfile, p, sbgfile(1:4) == 'none', rc)
         if (rc /= 0) then
            ierr = -1
            return
         endif
      else
         ! regular grid: mask on the full nmax x mmax grid (n fastest)
         allocate(mskg(nmax, mmax))
         mskg = 0
         if (inputformat(1:3) == 'asc') then
            open(newunit=u, file=trim(mskfile), status='old', action='read')
            allocate(tmp(mmax))
            do n = 1, nmax
               read(u, *) tmp
               do m = 1, mmax
                  mskg(n, m) = int(tmp(m), c_int8_t)
               enddo
            enddo
            close(u)
            deallocate(tmp)
         else
            open(newunit=u, file=trim(indexfile), form='unformatted', access='stream', status='old')
            read(u) np
            allocate(indices(np), msk(np))
            read(u) indices
            close(u)
            open(newunit=u, file=trim(mskfile), form='unformatted', access='stream', status='old')
            read(u) msk
            close(u)
            do i = 1, np
               m = (indices(i) - 1) / nmax + 1
               n = indices(i) - (m - 1) * nmax
               mskg(n, m) = msk(i)
            enddo
            deallocate(msk)
         endif
         rc = sfx_model_create_regular(p, int(nmax, c_int32_t), int(mmax, c_int32_t), x0, y0, dx, dy, rot, &
                                       c_loc(mskg), model)
         if (rc < 0) then
            call fail('mesh')
            ierr = -1
            return
         endif
         call table_size('kcs', 1, host_np)
         allocate(zb(host_np))
         if (sbgfile(1:4) /= 'none') then
            ! subgrid tables replace the bed level (initialize_bathymetry, sfincs_domain.f90:1588-1640)
         else if (inputformat(1:3) == 'asc') then
            allocate(zbg(nmax, mmax))
            open(newunit=u, file=trim(depfile), status='old', action='read')
            do n = 1, nmax
               read(u, *) (zbg(n, m), m = 1, mmax)
            enddo
            close(u)
            j = 0
            do m = 1, mmax
               do n = 1, nmax
                  if (mskg(n, m) > 0) then
                     j = j + 1
                     zb(j) = zbg(n, m)
                  endif
               enddo
            enddo
         else
            open(newunit=u, file=trim(depfile), form='unformatted', access='stream', status='old')
            read(u) zb
            close(u)
         endif
         if (sbgfile(1:4) == 'none') then
            rc = sfx_model_set_bathymetry(model, c_loc(zb))
            if (rc < 0) then
               call fail('bathymetry')
               ierr = -1
               return
            endif
         endif
      endif
      ! subgrid tables (read_subgrid_file, sfincs_subgrid.F90:20-53): a file that opens as
      ! NetCDF is read as such (huthresh 0), any other as the legacy binary format
      if (sbgfile(1:4) /= 'none' .and. is_netcdf(sbgfile)) then
         call logmsg('Info   : reading sub-grid netCDF file')
         rc = sfx_model_read_subgrid_netcdf(model, trim(sbgfile)//c_null_char)
         if (rc < 0) then
            call fail('subgrid file')
            ierr = -1
            return
         endif
      elseif (sbgfile(1:4) /= 'none') then
         call logmsg('Info    : reading subgrid file '//trim(sbgfile))
         rc = sfx_model_read_subgrid_binary(model, trim(sbgfile)//c_null_char)
         if (rc < 0) then
            call fail('subgrid file')
            ierr = -1
            return
         endif
      endif

      ! water level boundaries (sfincs_boundaries.f90:8-452)
      if (bndfile(1:4) /= 'none') then
         nb = count_lines(bndfile)
         allocate(xb(nb), yb(nb))
         open(newunit=u, file=trim(bndfile), status='old', action='read')
         do i = 1, nb
            read(u, *) xb(i), yb(i)
         enddo
         close(u)
         if (bzsfile(1:4) /= 'none') then
            nt = count_lines(bzsfile)
            allocate(tb(nt), zsb(nb, nt))
            open(newunit=u, file=trim(bzsfile), status='old', action='read')
            do i = 1, nt
               read(u, *) tb(i), (zsb(j, i), j = 1, nb)
            enddo
            close(u)
            rc = sfx_model_set_boundary(model, int(nb, c_int32_t), c_loc(xb), c_loc(yb), int(nt, c_int32_t), &
                                        c_loc(tb), c_loc(zsb))
            ! incoming infragravity levels, one row per bzs time (sfincs_boundaries.f90:126-140)
            if (rc >= 0 .and. bzifile(1:4) /= 'none') then
               allocate(zsi(nb, nt))
               open(newunit=u, file=trim(bzifile), status='old', action='read')
               do i = 1, nt
                  read(u, *) tdum, (zsi(j, i), j = 1, nb)
               enddo
               close(u)
               rc = sfx_model_set_boundary_ig(model, c_loc(zsi))
               deallocate(zsi)
            endif
         else
            rc = sfx_model_set_boundary(model, int(nb, c_int32_t), c_loc(xb), c_loc(yb), 0_c_int32_t, &
                                        c_null_ptr, c_null_ptr)
         endif
         if (rc < 0) call fail('boundaries')
      endif
      ! point discharges (sfincs_discharges.f90:8-320)
      if (srcfile(1:4) /= 'none') then
         nb = count_lines(srcfile)
         nt = count_lines(disfile)
         if (allocated(xb)) deallocate(xb, yb)
         allocate(xb(nb), yb(nb), tt(nt))
         if (allocated(zsb)) deallocate(zsb)
         allocate(zsb(nb, nt))
         open(newunit=u, file=trim(srcfile), status='old', action='read')
         do i = 1, nb
            read(u, *) xb(i), yb(i)
         enddo
         close(u)
         open(newunit=u, file=trim(disfile), status='old', action='read')
         do i = 1, nt
            read(u, *) tt(i), (zsb(j, i), j = 1, nb)
         enddo
         close(u)
         rc = sfx_model_set_sources(model, int(nb, c_int32_t), c_loc(xb), c_loc(yb), int(nt, c_int32_t), &
                                    c_loc(tt), c_loc(zsb))
         if (rc < 0) call fail('discharges')
         deallocate(tt)
      endif
      ! drainage structures: x_in y_in x_out y_out type and 1 (types 1-3) or 6 (gates)
      ! parameters per line (sfincs_discharges.f90:176-224)
      if (drnfile(1:4) /= 'none') then
         nb = count_lines(drnfile)
         if (allocated(xb)) deallocate(xb, yb)
         if (allocated(tt)) deallocate(tt)
         if (allocated(v1)) deallocate(v1)
         allocate(xb(nb), yb(nb), tt(nb), v1(nb), drn_type(nb), drn_par(6, nb))
         drn_par = 0.0
         open(newunit=u, file=trim(drnfile), status='old', action='read')
         do i = 1, nb
            read(u, '(a)') line
            read(line, *) xb(i), yb(i), tt(i), v1(i), drn_type(i)
            if (drn_type(i) >= 4) then
               read(line, *, iostat=ios) xb(i), yb(i), tt(i), v1(i), drn_type(i), (drn_par(j, i), j = 1, 6)
            else
               read(line, *, iostat=ios) xb(i), yb(i), tt(i), v1(i), drn_type(i), drn_par(1, i)
            endif
            if (ios /= 0) then
               call logmsg('Error   : drainage structure line with too few parameters: '//trim(line))
               ierr = -1
               return
            endif
         enddo
         close(u)
         rc = sfx_model_set_drainage(model, int(nb, c_int32_t), c_loc(xb), c_loc(yb), c_loc(tt), c_loc(v1), &
                                     c_loc(drn_type), c_loc(drn_par))
         if (rc < 0) then
            call fail('drainage structures')
            ierr = -1
            return
         endif
         deallocate(tt, v1, drn_type, drn_par)
      endif
      ! gridded meteo files, read after the mesh (read_meteo_data, sfincs_meteo.f90:5-245)
      do j = 1, 5
         if (metfile(j)(1:4) /= 'none') then
            rc = sfx_model_read_meteo_file(model, trim(metkind(j))//c_null_char, trim(metfile(j))//c_null_char, &
                                           trim(tref)//c_null_char)
            if (rc < 0) then
               call fail(trim(metkey(j)))
               ierr = -1
               return
            endif
         endif
      enddo
      ! observation points: x y ['name'] per line (read_obs_points, sfincs_obspoints.f90:24-81)
      nobs = 0
      if (allocated(obs_nm)) deallocate(obs_nm, obs_x, obs_y)
      if (allocated(his_tp)) deallocate(his_tp, his_vp)
      if (obsfile(1:4) /= 'none') then
         open(newunit=u, file=trim(obsfile), status='old', action='read', iostat=ios)
         if (ios /= 0) then
            call logmsg('Error   : cannot read observation points file '//trim(obsfile))
            ierr = -1
            return
         endif
         do
            read(u, *, iostat=ios) rdum
            if (ios < 0) exit
            nobs = nobs + 1
         enddo
         rewind(u)
         allocate(obs_nm(nobs), obs_x(nobs), obs_y(nobs))
         do i = 1, nobs
            read(u, '(a)') line
            read(line, *) obs_x(i), obs_y(i)
         enddo
         close(u)
         rc = sfx_model_find_cells(model, int(nobs, c_int32_t), c_loc(obs_x), c_loc(obs_y), c_loc(obs_nm))
         do i = 1, nobs
            if (obs_nm(i) == 0) then
               write(line, '(a,i0,a)') 'Warning ! Observation point ', i, ' falls outside model domain !'
               call logmsg(trim(line))
            endif
         enddo
      endif
      ! spatially varying constant infiltration: float32 mm/h per active cell (sfincs_infiltration.f90:241-249)
      if (qinffile(1:4) /= 'none') then
         if (allocated(tmp)) deallocate(tmp)
         allocate(tmp(host_np))
         open(newunit=u, file=trim(qinffile), form='unformatted', access='stream', status='old', iostat=ios)
         if (ios == 0) read(u, iostat=ios) tmp
         if (ios /= 0) then
            call logmsg('Error   : cannot read infiltration file '//trim(qinffile))
            ierr = -1
            return
         endif
         close(u)
         rc = sfx_model_set_infiltration(model, c_loc(tmp))
         if (rc < 0) then
            call fail('infiltration')
            ierr = -1
            return
         endif
         deallocate(tmp)
      endif
      ! curve-number retention S, float32 inches per active cell (sfincs_infiltration.f90:277-286)
      if (scsfile(1:4) /= 'none') then
         if (allocated(tmp)) deallocate(tmp)
         allocate(tmp(host_np))
         open(newunit=u, file=trim(scsfile), form='unformatted', access='stream', status='old', iostat=ios)
         if (ios == 0) read(u, iostat=ios) tmp
         if (ios /= 0) then
            call logmsg('Error   : cannot read scs file '//trim(scsfile))
            ierr = -1
            return
         endif
         close(u)
         rc = sfx_model_set_curve_number(model, c_loc(tmp))
         if (rc < 0) then
            call fail('curve number')
            ierr = -1
            return
         endif
         deallocate(tmp)
      endif
      ! thin dams and weirs (read_structures, sfincs_structures.f90:8-35): polylines
      if (thdfile(1:4) /= 'none') then
         call read_polylines(thdfile, 0, pl_nv, pl_x, pl_y, pl_p, pl_ok)
         if (.not. pl_ok) then
            call logmsg('Error   : cannot read thin dam file '//trim(thdfile))
            ierr = -1
            return
         endif
         rc = sfx_model_set_thin_dams(model, int(size(pl_nv), c_int32_t), c_loc(pl_nv), c_loc(pl_x), c_loc(pl_y))
         if (rc < 0) then
            call fail('thin dams')
            ierr = -1
            return
         endif
         deallocate(pl_nv, pl_x, pl_y, pl_p)
      endif
      if (weirfile(1:4) /= 'none') then
         call read_polylines(weirfile, 2, pl_nv, pl_x, pl_y, pl_p, pl_ok)
         if (.not. pl_ok) then
            call logmsg('Error   : cannot read weir file '//trim(weirfile))
            ierr = -1
            return
         endif
         rc = sfx_model_set_weirs(model, int(size(pl_nv), c_int32_t), c_loc(pl_nv), c_loc(pl_x), c_loc(pl_y), &
                                  c_loc(pl_p))
         if (rc < 0) then
            call fail('weirs')
            ierr = -1
            return
         endif
         deallocate(pl_nv, pl_x, pl_y, pl_p)
      endif
      ! downstream river points: x y x_in y_in slope distance (sfincs_boundaries.f90:221-268)
      if (bdrfile(1:4) /= 'none') then
         nb = count_lines(bdrfile)
         if (allocated(tt)) deallocate(tt)
         if (allocated(v1)) deallocate(v1)
         if (allocated(v2)) deallocate(v2)
         if (allocated(tmp)) deallocate(tmp)
         allocate(tt(nb), v1(nb), v2(nb), tmp(nb))
         if (allocated(xb)) deallocate(xb, yb)
         allocate(xb(nb), yb(nb))
         open(newunit=u, file=trim(bdrfile), status='old', action='read')
         do i = 1, nb
            read(u, *) xb(i), yb(i), tt(i), v1(i), v2(i), tmp(i)
         enddo
         close(u)
         rc = sfx_model_set_river_boundary(model, int(nb, c_int32_t), c_loc(xb), c_loc(yb), c_loc(tt), c_loc(v1), &
                                           c_loc(v2), c_loc(tmp))
         if (rc < 0) call fail('river boundaries')
         deallocate(tt, v1, v2, tmp)
      endif
      ! uniform rain and wind series (sfincs_meteo.f90:247-312)
      if (prcfile(1:4) /= 'none') then
         nt = count_lines(prcfile)
         allocate(tt(nt), v1(nt))
         open(newunit=u, file=trim(prcfile), status='old', action='read')
         do i = 1, nt
            read(u, *) tt(i), v1(i)
         enddo
         close(u)
         rc = sfx_model_set_rain(model, int(nt, c_int32_t), c_loc(tt), c_loc(v1))
         if (rc < 0) call fail('rain')
         if (allocated(his_tp)) deallocate(his_tp, his_vp)
         allocate(his_tp(nt), his_vp(nt))
         his_tp = tt
         his_vp = v1
         deallocate(tt, v1)
      endif
      if (wndfile(1:4) /= 'none') then
         nt = count_lines(wndfile)
         allocate(tt(nt), v1(nt), v2(nt))
         open(newunit=u, file=trim(wndfile), status='old', action='read')
         do i = 1, nt
            read(u, *) tt(i), v1(i), v2(i)
         enddo
         close(u)
         rc = sfx_model_set_wind(model, int(nt, c_int32_t), c_loc(tt), c_loc(v1), c_loc(v2))
         if (rc < 0) call fail('wind')
         deallocate(tt, v1, v2)
      endif
      if (rstfile(1:4) /= 'none') then   ! binary restart (sfincs_initial_conditions.F90:44-54)
         call logmsg('Info    : reading restart file '//trim(rstfile))
         rc = sfx_model_set_restart(model, trim(rstfile)//c_null_char)
         if (rc < 0) then
            call fail('restart file')
            ierr = -1
            return
         endif
      endif
      rc = sfx_model_initialize(model)
      if (rc < 0) then
         call fail('initialize')
         ierr = -1
         return
      endif
      nbytes = int(c_sizeof(p), c_size_t)
      rc = sfx_model_get_table(model, 'params'//c_null_char, c_loc(p), nbytes)
      host_huthresh = p%huthresh           ! 0 once NetCDF-style subgrid tables are set
      host_subgrid = sbgfile(1:4) /= 'none'
      out_zvolume = out_zvolume .and. host_subgrid
      ierr = 0
   end function

   type(c_ptr) function host_model()
      host_model = model
   end function

   ! The library's last error message, NUL-terminated, into buf(1:n).
   subroutine host_last_error(buf, n)
      integer, intent(in) :: n
      character(kind=c_char), intent(out) :: buf(n)
      integer(c_int) :: rc
      buf = c_null_char
      rc = sfx_last_error(buf, int(n, c_size_t))
   end subroutine

   ! The active cell (1-based, 0 outside) holding (x, y).
   integer function host_find_cell(x, y)
      real(c_double), intent(in) :: x, y
      integer(c_int32_t) :: idx
      integer(c_int) :: rc
      idx = 0
      if (c_associated(model)) rc = sfx_model_find_cell(model, real(x, c_float), real(y, c_float), idx)
      host_find_cell = idx
   end function

   ! set_logical("qext") (sfincs_bmi.f90:226-245): external sources on or off; off writes
   ! zeros to the device, which adds nothing. -1 when the model has no qext field.
   integer function host_qext_switch(on)
      logical, intent(in) :: on
      real*4, allocatable, target :: z(:)
      integer(c_int) :: rc
      host_qext_switch = -1
      if (.not. allocated(host_qext)) return
      host_qext_on = on
      if (.not. on .and. c_associated(handle)) then
         allocate(z(host_np))
         z = 0.0
         rc = sfx_set_state(handle, 'qext'//c_null_char, c_loc(z), int(4 * host_np, c_size_t))
         deallocate(z)
      endif
      host_qext_switch = 0
   end function

   ! Host copies of the mesh fields the BMI hands out (get_value_ptr, get_sfincs_cell_area).
   subroutine host_mesh_fields()
      integer :: n
      call table_size('z_xz', 4, n)
      if (allocated(host_xz)) deallocate(host_xz, host_yz, host_zb, host_zmin, host_iref, host_area)
      host_np = n
      allocate(host_xz(n), host_yz(n), host_iref(n))
      call get_f32('z_xz', host_xz); call get_f32('z_yz', host_yz)
      call table_size('zb', 4, n); allocate(host_zb(n)); if (n > 0) call get_f32('zb', host_zb)
      call table_size('z_zmin', 4, n); allocate(host_zmin(n)); if (n > 0) call get_f32('z_zmin', host_zmin)
      call table_size('cell_area', 4, n); allocate(host_area(n)); call get_f32('cell_area', host_area)
      n = sfx_model_get_table(model, 'z_flags_iref'//c_null_char, c_loc(host_iref), int(size(host_iref), c_size_t))
      if (allocated(host_zn)) deallocate(host_zn, host_zm)
      allocate(host_zn(host_np), host_zm(host_np))
      call get_i32('z_index_z_n', host_zn); call get_i32('z_index_z_m', host_zm)
   contains
      subroutine get_i32(name, a)
         character(*), intent(in) :: name
         integer(c_int32_t), target, intent(inout) :: a(:)
         integer :: rc
         rc = sfx_model_get_table(model, trim(name)//c_null_char, c_loc(a), int(4 * size(a), c_size_t))
      end subroutine
      subroutine get_f32(name, a)
         character(*), intent(in) :: name
         real*4, target, intent(inout) :: a(:)
         integer :: rc
         rc = sfx_model_get_table(model, trim(name)//c_null_char, c_loc(a), int(4 * size(a), c_size_t))
      end subroutine
   end subroutine

   subroutine table_size(name, esize, n)
      character(*), intent(in) :: name
      integer, intent(in) :: esize
      integer, intent(out) :: n
      integer(c_size_t) :: nb
      integer :: rc
      rc = sfx_model_table_bytes(model, trim(name)//c_null_char, nb)
      n = int(nb / esize)
   end subroutine

   subroutine read_quadtree(qtrfile, mskfile, p, set_zb, rc)
      ! legacy binary quadtree file (sfincs_quadtree.F90:210-292) + binary mask file; the
      ! bed level from the file unless subgrid tables replace it
      character(256), intent(in) :: qtrfile, mskfile
      type(sfx_params), intent(in) :: p
      logical, intent(in) :: set_zb
      integer, intent(out) :: rc
      type(sfx_quadtree_input) :: q
      integer(c_int8_t) :: iversion, nlev
      integer(c_int32_t) :: iepsg, np
      real(c_float) :: x0, y0, dx, dy, rot
      integer(c_int8_t), allocatable, target, save :: lev(:), fnu(:), fmu(:), fnd(:), fmd(:), msk(:)
      integer(c_int32_t), allocatable, target, save :: n(:), m(:), nu1(:), nu2(:), mu1(:), mu2(:)
      integer(c_int32_t), allocatable, target, save :: nd1(:), nd2(:), md1(:), md2(:)
      real(c_float), allocatable, target, save :: zz(:), zb(:)
      integer :: u, i, j
      open(newunit=u, file=trim(qtrfile), form='unformatted', access='stream', status='old')
      read(u) iversion, iepsg, np, nlev, x0, y0, dx, dy, rot
      allocate(lev(np), n(np), m(np), fnu(np), nu1(np), nu2(np), fmu(np), mu1(np), mu2(np))
      allocate(fnd(np), nd1(np), nd2(np), fmd(np), md1(np), md2(np), zz(np), msk(np))
      read(u) lev, n, m, fnu, nu1, nu2, fmu, mu1, mu2, fnd, nd1, nd2, fmd, md1, md2, zz
      close(u)
      open(newunit=u, file=trim(mskfile), form='unformatted', access='stream', status='old')
      read(u) msk
      close(u)
      q%nr_points = np; q%nr_levels = nlev; q%x0 = x0; q%y0 = y0; q%dx = dx; q%dy = dy; q%rotation_deg = rot
      q%level = c_loc(lev); q%md = c_loc(fmd); q%mu = c_loc(fmu); q%nd = c_loc(fnd); q%nu = c_loc(fnu)
      q%mask = c_loc(msk); q%n = c_loc(n); q%m = c_loc(m); q%md1 = c_loc(md1); q%md2 = c_loc(md2)
      q%mu1 = c_loc(mu1); q%mu2 = c_loc(mu2); q%nd1 = c_loc(nd1); q%nd2 = c_loc(nd2)
      q%nu1 = c_loc(nu1); q%nu2 = c_loc(nu2)
      rc = sfx_model_create_quadtree(p, q, model)
      if (rc < 0) then
         call fail('quadtree mesh')
         return
      endif
      if (.not. set_zb) return
      ! bed level from the quadtree file for active points (sfincs_domain.f90:1635-1637)
      allocate(zb(count(msk > 0)))
      j = 0
      do i = 1, np
         if (msk(i) > 0) then
            j = j + 1
            zb(j) = zz(i)
         endif
      enddo
      rc = sfx_model_set_bathymetry(model, c_loc(zb))
      if (rc < 0) call fail('bathymetry')
      rc = min(rc, 0)
   end subroutine

   ! ---- sfincs_update (sfincs_lib.f90:337-683) ----------------------------------------
   function sfincs_update(dtrange) result(ierr)
      double precision, intent(in) :: dtrange
      integer :: ierr
      type(sfx_step_stats) :: st
      integer(c_int) :: rc
      real*8 :: tend, tnext
      if (host_qext_on) then
         rc = sfx_set_state(handle, 'qext'//c_null_char, c_loc(host_qext), int(4 * host_np, c_size_t))
         if (rc < 0) then
            call fail('qext upload')
            ierr = -1
            return
         endif
      endif
      ierr = host_upload_exposed()
      if (ierr /= 0) return
      if (dtrange < -999.0d0 .or. dtrange > 0.0d0) then
         ! run to the end of the interval, stopping at each output time on the way: the
         ! device stops at the first step whose time reaches it, the step at which the
         ! reference's trigger fires
         tend = host_t1
         if (dtrange > 0.0d0) tend = t + dtrange
         rc = 0
         do while (t < tend)
            tnext = min(tend, tmapout, tmaxout, trstout, thisout)
            if (tnext <= t) then
               rc = sfx_update(handle, 0.0d0, st)      ! an output is due: one step
            else
               rc = sfx_update_to(handle, real(tnext, c_double), st)
            endif
            if (rc < 0) exit
            call after_steps(st, rc)
            if (rc /= 0) exit
         enddo
      else
         rc = sfx_update(handle, 0.0d0, st)             ! one step (BMI update)
         if (rc >= 0) call after_steps(st, rc)
      endif
      ierr = 0
      if (host_zs_out) call host_sync()
      if (rc < 0) then
         call fail('update')
         ierr = -1
      else if (rc == 1 .or. st%error /= 0) then
         call logmsg('Error! Minimum time step reached. Simulation stopped.')
         ierr = 1
      endif
   end function

   ! explicit device-to-host copy of zs and zsmax (the form of `acc update host`)
   subroutine host_sync()
      integer(c_int) :: rc
      rc = sfx_get_state(handle, 'zs'//c_null_char, c_loc(host_zs), int(8 * host_np, c_size_t))
      rc = sfx_get_state(handle, 'zsmax'//c_null_char, c_loc(host_zsmax), int(4 * host_np, c_size_t))
      if (host_zs_out) shadow_zs = host_zs
   end subroutine

   ! get_value_ptr hands out zs or zb: from then on their host copies are live for the coupler
   subroutine host_expose(name)
      character(*), intent(in) :: name
      select case (name)
      case ('zs')
         host_zs_out = .true.
         call host_sync()
      case ('zb')
         if (.not. host_zb_out .and. size(host_zb) > 0) then
            host_zb_out = .true.
            shadow_zb = host_zb
         endif
      end select
   end subroutine

   ! Uploads the copies a coupler changed since they were handed out or last uploaded.
   integer function host_upload_exposed() result(ierr)
      integer(c_int) :: rc
      ierr = 0
      if (host_zs_out) then
         if (any(host_zs /= shadow_zs)) then
            rc = sfx_set_state(handle, 'zs'//c_null_char, c_loc(host_zs), int(8 * host_np, c_size_t))
            if (rc < 0) then
               call fail('zs upload')
               ierr = -1
               return
            endif
            shadow_zs = host_zs
         endif
      endif
      if (host_zb_out) then
         if (any(host_zb /= shadow_zb)) then
            rc = sfx_set_state(handle, 'zb'//c_null_char, c_loc(host_zb), int(4 * size(host_zb), c_size_t))
            if (rc < 0) then
               call fail('zb upload')
               ierr = -1
               return
            endif
            shadow_zb = host_zb
         endif
      endif
   end function

   ! update_zbuv (sfincs_bmi.f90:277-284): the bed written by the coupler reaches the device,
   ! then zbuvmx is derived again from it (compute_zbuvmx, sfincs_domain.f90:1746-1762)
   integer function host_update_zbuv() result(ierr)
      integer(c_int) :: rc
      ierr = -1
      if (.not. c_associated(handle)) return
      ierr = host_upload_exposed()
      if (ierr /= 0) return
      rc = sfx_update_zbuvmx(handle)
      if (rc < 0) then
         call fail('update_zbuv')
         ierr = -1
      endif
   end function

   ! A per-cell f32 state field of the device to a raw file (active cells, reference order).
   subroutine write_cell_f32(name, n, fname)
      character(*), intent(in) :: name, fname
      integer, intent(in) :: n
      real(c_float), allocatable, target :: buf(:)
      integer :: u
      integer(c_int) :: rc
      allocate(buf(n))
      rc = sfx_get_state(handle, name//c_null_char, c_loc(buf), int(4 * n, c_size_t))
      if (rc == 0) then
         open(newunit=u, file=fname, form='unformatted', access='stream', status='replace')
         write(u) buf
         close(u)
      endif
      deallocate(buf)
   end subroutine

   ! timestep_analysis_finalize and _write_log (sfincs_timestep_analysis.f90:77-247): the
   ! per-cell average required step and percentage limiting, and the most limiting uv-point.
   subroutine write_timestep_analysis()
      real(c_float), allocatable, target :: tsum(:), dtc(:), pct(:)
      integer(c_int32_t), allocatable, target :: twet(:), tlim(:)
      real(c_double) :: tt
      real(c_float) :: dtt
      integer(c_int64_t) :: nt
      integer(c_int) :: rc
      integer :: u, ip
      character(256) :: line
      allocate(tsum(host_npuv), twet(host_npuv), tlim(host_npuv), dtc(host_np), pct(host_np))
      rc = sfx_get_state(handle, 'tsa_sum'//c_null_char, c_loc(tsum), int(4 * host_npuv, c_size_t))
      if (rc == 0) rc = sfx_get_state(handle, 'tsa_wet'//c_null_char, c_loc(twet), int(4 * host_npuv, c_size_t))
      if (rc == 0) rc = sfx_get_state(handle, 'tsa_lim'//c_null_char, c_loc(tlim), int(4 * host_npuv, c_size_t))
      if (rc == 0) rc = sfx_get_time(handle, tt, dtt, nt)
      if (rc == 0 .and. nt > 0) rc = sfx_model_timestep_analysis_per_cell(model, c_loc(tsum), c_loc(twet), &
                                                                         c_loc(tlim), nt, c_loc(dtc), c_loc(pct))
      if (rc == 0 .and. nt > 0) then
         open(newunit=u, file='timestep_analysis_dt.dat', form='unformatted', access='stream', status='replace')
         write(u) dtc
         close(u)
         open(newunit=u, file='timestep_analysis_pct.dat', form='unformatted', access='stream', status='replace')
         write(u) pct
         close(u)
         call logmsg('----------- Time step analysis -----------')
         if (maxval(tlim) > 0) then
            ip = maxloc(tlim, dim=1)
            write(line, '(a,i12)') ' Index of most time step limiting uv point : ', ip
            call logmsg(trim(line))
            write(line, '(a,f10.1,a)') ' Percentage of time point was limiting     : ', &
                 real(maxval(tlim)) / real(max(maxval(twet), 1)) * 100.0, ' %'
            call logmsg(trim(line))
         endif
      endif
      deallocate(tsum, twet, tlim, dtc, pct)
   end subroutine

   ! ---- scheduled output ------------------------------------------------------------
   ! initialize_output (sfincs_output.f90:8-60) with the binary writers' files (:314-373)
   subroutine output_init()
      ntmapout = 0
      ntmaxout = 0
      call output_close()
      if (out_dtmap > 1.0d-6) then
         tmapout = out_t0
         open(newunit=u_zs, file='zs.dat', form='unformatted', status='replace')
         if (out_zvolume) open(newunit=u_zv, file='z_volume.dat', form='unformatted', status='replace')
      else
         tmapout = 1.0d9
      endif
      if (out_dtmax > 1.0d-6) then
         tmaxout = out_t0 + out_dtmax
         if (host_store_zsmax == 1) open(newunit=u_zsmax, file='zsmax.dat', form='unformatted', status='replace')
         if (host_flags(1) == 1) open(newunit=u_vmax, file='velmax.dat', form='unformatted', status='replace')
         if (host_flags(3) == 1) open(newunit=u_twet, file='tmax.dat', form='unformatted', status='replace')
         if (host_flags(2) == 1) open(newunit=u_qmax, file='qmax.dat', form='unformatted', status='replace')
      else
         tmaxout = 1.0d9
      endif
      nthisout = 0
      if (out_dthis > 1.0d-6 .and. nobs > 0) then   ! initialize_output (sfincs_output.f90:59-76)
         thisout = host_t0
         call his_delete('zst.txt')
         call his_delete('prcp.txt')
         call his_delete('qdrain.txt')
      else
         thisout = 1.0d9
      endif
      if (out_dtrst > 1.0d-6) then
         trstout = out_t0 + out_dtrst
      else if (out_trst > 1.0d-6) then
         trstout = out_trst
      else
         trstout = 1.0d9
      endif
   end subroutine

   subroutine output_close()
      if (u_zs /= 0) close(u_zs)
      if (u_zv /= 0) close(u_zv)
      if (u_zsmax /= 0) close(u_zsmax)
      if (u_vmax /= 0) close(u_vmax)
      if (u_twet /= 0) close(u_twet)
      if (u_qmax /= 0) close(u_qmax)
      u_zs = 0; u_zv = 0; u_zsmax = 0; u_vmax = 0; u_twet = 0; u_qmax = 0
   end subroutine

   ! After the device stopped: the output triggers of the step it stopped at
   ! (sfincs_lib.f90:419-489), the output (:624-630) and, at an instability stop, the last
   ! map and maximum output (:634-648).
   subroutine after_steps(st, rc)
      type(sfx_step_stats), intent(in) :: st
      integer(c_int), intent(in) :: rc
      logical :: wmap, wmax, wrst, whis
      real*8 :: tout
      t = st%t
      dt = st%dt
      if (rc == 1 .or. st%error /= 0) then
         ntmaxout = ntmaxout + 1
         call write_output(t, u_zs /= 0, .true., .false.)
         if (thisout < 1.0d8) call write_his(t)
         return
      endif
      wmap = .false.; wmax = .false.; wrst = .false.; whis = .false.
      tout = t
      if (t >= tmapout) then
         wmap = .true.
         ntmapout = ntmapout + 1
         tout = max(tmapout, t - dt)
         tmapout = tmapout + out_dtmap
      endif
      if (t >= tmaxout) then
         wmax = .true.
         ntmaxout = ntmaxout + 1
         tout = max(tmaxout, t - dt)
         if (t < host_t1) tmaxout = tmaxout + out_dtmax
      endif
      if (t >= trstout) then
         wrst = .true.
         tout = trstout
         if (out_dtrst > 1.0d-6) then
            trstout = trstout + out_dtrst
         else
            trstout = 1.0d9
         endif
      endif
      if (t >= thisout) then
         whis = .true.
         nthisout = nthisout + 1
         tout = max(thisout, t - dt)
         thisout = thisout + out_dthis
      endif
      if (wmap .or. wmax .or. wrst) call write_output(tout, wmap, wmax, wrst)
      if (whis) call write_his(tout)
   end subroutine

   subroutine his_delete(name)
      character(*), intent(in) :: name
      integer :: u, ios
      open(newunit=u, file=name, iostat=ios)
      if (ios == 0) close(u, status='delete')
   end subroutine

   ! write_his_output (sfincs_output.f90:593-655), with the binary-output reference's text
   ! files: the level at each observation point (-999 outside the mesh), the rain of the
   ! step's series there in mm/h, the drainage structures' intake discharges
   subroutine write_his(tout)
      real*8, intent(in) :: tout
      real*4, allocatable :: zobs(:), tprcp(:)
      real(c_float), allocatable, target :: qs(:)
      real*4 :: twfac, ptmp
      integer :: i, u, itp, nt
      integer(c_int) :: r
      if (nobs == 0) return
      call host_sync()
      allocate(zobs(nobs), tprcp(nobs))
      ptmp = 0.0
      if (allocated(his_tp)) then   ! update_precipitation_from_timeseries at the step's time
         nt = size(his_tp)
         do itp = 1, nt
            if (his_tp(itp) > t) exit
         enddo
         itp = max(min(itp, nt), 2)
         twfac = real((t - dble(his_tp(itp - 1))) / dble(his_tp(itp) - his_tp(itp - 1)), 4)
         ptmp = (his_vp(itp - 1) * (1.0 - twfac) + his_vp(itp) * twfac) / (1000 * 3600)
      endif
      do i = 1, nobs
         if (obs_nm(i) > 0) then
            zobs(i) = real(host_zs(obs_nm(i)), 4)
            tprcp(i) = ptmp * 1000 * 3600
         else
            zobs(i) = -999.0
            tprcp(i) = -999.0
         endif
      enddo
      open(newunit=u, file='zst.txt', position='append', status='unknown')
      write(u, '(f12.1,10000f9.3)') tout, zobs
      close(u)
      if (allocated(his_tp)) then
         open(newunit=u, file='prcp.txt', position='append', status='unknown')
         write(u, '(f12.1,10000f9.3)') tout, tprcp
         close(u)
      endif
      if (his_ndrn > 0 .and. his_qdrain) then
         allocate(qs(his_nsrc + 2 * his_ndrn))
         r = sfx_get_state(handle, 'qtsrc'//c_null_char, c_loc(qs), int(4 * size(qs), c_size_t))
         open(newunit=u, file='qdrain.txt', position='append', status='unknown')
         write(u, '(f12.1,10000f9.3)') tout, (qs(i), i = his_nsrc + 1, his_nsrc + 2 * his_ndrn, 2)
         close(u)
      endif
   end subroutine

   ! write_output (sfincs_output.f90:80-258), binary flavour: the fields come from the
   ! device first (the form of `acc update host`); after a maximum output the maxima start
   ! again from -999 (twet from 0) on the device, as the reference resets them
   subroutine write_output(tout, wmap, wmax, wrst)
      real*8, intent(in) :: tout
      logical, intent(in) :: wmap, wmax, wrst
      real*8, allocatable, target :: v8(:)
      real*4, allocatable, target :: f4(:)
      integer(c_int) :: rc
      integer :: nm
      character(15) :: ts
      if (wmap .and. u_zs /= 0) then
         call host_sync()
         write(u_zs) host_zs
         if (u_zv /= 0) then
            allocate(v8(host_np))
            rc = sfx_get_state(handle, 'z_volume'//c_null_char, c_loc(v8), int(8 * host_np, c_size_t))
            write(u_zv) v8
            deallocate(v8)
         endif
      endif
      if (wmax .and. out_dtmax > 0.0d0) then
         allocate(f4(host_np))
         if (u_zsmax /= 0) then
            rc = sfx_get_state(handle, 'zsmax'//c_null_char, c_loc(f4), int(4 * host_np, c_size_t))
            ! dry cells: -999 (write_max_output, sfincs_output.f90:466-487)
            if (host_subgrid) then
               do nm = 1, host_np
                  if (f4(nm) < host_zmin(nm) + host_huthresh) f4(nm) = -999.0
               enddo
            else
               do nm = 1, host_np
                  if (f4(nm) < host_zb(nm) + host_huthresh) f4(nm) = -999.0
               enddo
            endif
            write(u_zsmax) f4
            f4 = -999.0
            rc = sfx_set_state(handle, 'zsmax'//c_null_char, c_loc(f4), int(4 * host_np, c_size_t))
         endif
         if (u_vmax /= 0) call max_field('vmax', u_vmax, -999.0)
         if (u_twet /= 0) call max_field('twet', u_twet, 0.0)
         if (u_qmax /= 0) call max_field('qmax', u_qmax, -999.0)
         deallocate(f4)
      endif
      if (wrst) then
         ts = time_string(tout)
         call logmsg('Info    : writing restart file sfincs.'//ts//'.rst')
         rc = sfx_write_restart(handle, 'sfincs.'//ts//'.rst'//c_null_char)
         if (rc < 0) call fail('restart file')
      endif
   contains
      subroutine max_field(name, u, reset)
         character(*), intent(in) :: name
         integer, intent(in) :: u
         real*4, intent(in) :: reset
         integer(c_int) :: r
         r = sfx_get_state(handle, name//c_null_char, c_loc(f4), int(4 * host_np, c_size_t))
         write(u) f4
         f4 = reset
         r = sfx_set_state(handle, name//c_null_char, c_loc(f4), int(4 * host_np, c_size_t))
      end subroutine
   end subroutine

   ! finalize_output (sfincs_output.f90:262-310): the maxima of the last (or only) interval
   subroutine output_finalize()
      if (out_dtmax > 1.0d-6 .and. ntmaxout == 0) then
         ntmaxout = 1
         call write_output(t, .false., .true., .false.)
      else if (out_dtmax > 1.0d-6 .and. ntmaxout > 0 .and. t < tmaxout) then
         ntmaxout = ntmaxout + 1
         call write_output(host_t1, .false., .true., .false.)
      endif
      call output_close()
   end subroutine

   ! time_to_string (sfincs_date.f90:298-327): 'yyyymmdd.hhmmss' of int(t) seconds after tref
   character(15) function time_string(tsec)
      real*8, intent(in) :: tsec
      integer(8) :: total, days, secs, l, n, i, j, y, mo, d
      total = int(host_tref_sec, 8) + int(tsec, 8)
      days = total / 86400
      secs = total - days * 86400
      ! Julian day number to the proleptic Gregorian date (inverse of date_seconds)
      l = days + 68569
      n = 4 * l / 146097
      l = l - (146097 * n + 3) / 4
      i = 4000 * (l + 1) / 1461001
      l = l - 1461 * i / 4 + 31
      j = 80 * l / 2447
      d = l - 2447 * j / 80
      l = j / 11
      mo = j + 2 - 12 * l
      y = 100 * (n - 49) + i + l
      write(time_string, '(i4.4,i2.2,i2.2,a1,i2.2,i2.2,i2.2)') y, mo, d, '.', secs / 3600, mod(secs, 3600_8) / 60, &
         mod(secs, 60_8)
   end function

   ! ---- sfincs_finalize (sfincs_lib.f90:687-819) ---------------------------------------
   function sfincs_finalize() result(ierr)
      integer :: ierr, u
      integer(c_int) :: rc
      ierr = 0
      if (c_associated(handle)) then
         if (host_flags(4) == 1) call write_timestep_analysis()
         call output_finalize()
         call host_sync()
         open(newunit=u, file='zs_final.dat', form='unformatted', access='stream', status='replace')
         write(u) host_zs
         close(u)
         rc = sfx_destroy(handle)
         handle = c_null_ptr
      endif
      if (c_associated(model)) then
         call sfx_model_destroy(model)
         model = c_null_ptr
      endif
      call logmsg('---------- Simulation finished -----------')
      if (logu /= 0) close(logu)
      logu = 0
   end function

end module sfincs_lib


! BMI / XMI names of the reference's shared library (sfincs_bmi.f90:56-387, subset).
module sfincs_bmi
   use iso_c_binding
   use sfincs_lib
   implicit none
   ! string lengths of the BMI (sfincs_bmi.f90:41-52)
   integer(c_int), bind(C, name='BMI_LENVARADDRESS') :: BMI_LENVARADDRESS = 64
   integer(c_int), bind(C, name='BMI_LENVARTYPE') :: BMI_LENVARTYPE = 64
   integer(c_int), bind(C, name='BMI_LENGRIDTYPE') :: BMI_LENGRIDTYPE = 64
   integer(c_int), bind(C, name='BMI_LENCOMPONENTNAME') :: BMI_LENCOMPONENTNAME = 64
   integer(c_int), bind(C, name='BMI_LENVERSION') :: BMI_LENVERSION = 256
   integer(c_int), bind(C, name='BMI_LENERRMESSAGE') :: BMI_LENERRMESSAGE = 1024
contains
   ! initialize() (sfincs_bmi.f90:56-63): reads sfincs.inp of the working directory
   integer(c_int) function initialize() result(ierr) bind(C, name='initialize')
      host_qext_requested = .true.   ! the BMI always carries the qext field (sfincs_bmi.f90:60)
      ierr = sfincs_initialize()
   end function
   integer(c_int) function update() result(ierr) bind(C, name='update')
      ierr = sfincs_update(-1.0d0)
   end function
   integer(c_int) function update_until(tt) result(ierr) bind(C, name='update_until')
      real(c_double), value :: tt
      ierr = sfincs_update(tt - t)
   end function
   integer(c_int) function finalize() result(ierr) bind(C, name='finalize')
      ierr = sfincs_finalize()
   end function
   ! Deck reader alone (no GPU): builds the host model that initialize would move to HBM.
   integer(c_int) function sfincs_host_read_model() result(ierr) bind(C, name='sfincs_host_read_model')
      ierr = sfincs_read_model()
      if (ierr == 0) call host_mesh_fields()
   end function
   type(c_ptr) function sfincs_host_model() bind(C, name='sfincs_host_model')
      sfincs_host_model = host_model()
   end function
   integer(c_int) function get_start_time(tt) result(ierr) bind(C, name='get_start_time')
      real(c_double), intent(out) :: tt
      tt = host_t0
      ierr = 0
   end function
   integer(c_int) function get_end_time(tt) result(ierr) bind(C, name='get_end_time')
      real(c_double), intent(out) :: tt
      tt = host_t1
      ierr = 0
   end function
   integer(c_int) function get_current_time(tt) result(ierr) bind(C, name='get_current_time')
      real(c_double), intent(out) :: tt
      tt = t
      ierr = 0
   end function
   integer(c_int) function get_time_step(tt) result(ierr) bind(C, name='get_time_step')
      real(c_double), intent(out) :: tt
      tt = dt
      ierr = 0
   end function
   ! Host pointers to zs / zsmax / zb after an explicit device-to-host copy
   ! (the reference hands out c_loc of its arrays, sfincs_bmi.f90:95-129).
   ! get_value_ptr (sfincs_bmi.f90:95-129): zs and zsmax after an explicit device-to-host
   ! copy; the mesh fields z_xz, z_yz, zb, subgrid_z_zmin from their host copies. qext and
   ! uorb (coupling inputs of processes outside this build) are not available.
   integer(c_int) function get_value_ptr(c_var_name, x) result(ierr) bind(C, name='get_value_ptr')
      character(kind=c_char), intent(in) :: c_var_name(*)
      type(c_ptr), intent(out) :: x
      character(64) :: nm
      nm = bmi_name(c_var_name)
      ierr = 0
      select case (trim(nm))
      case ('zs')
         call host_expose('zs')
         x = c_loc(host_zs)
      case ('zsmax')
         call host_sync()
         x = c_loc(host_zsmax)
      case ('z_xz')
         x = c_loc(host_xz)
      case ('z_yz')
         x = c_loc(host_yz)
      case ('zb')
         call host_expose('zb')
         x = c_loc(host_zb)
      case ('subgrid_z_zmin')
         x = c_loc(host_zmin)
      case ('qext')
         if (allocated(host_qext)) then
            x = c_loc(host_qext)
         else
            x = c_null_ptr
            ierr = -1
         endif
      case default
         x = c_null_ptr
         ierr = -1
      end select
   end function

   character(64) function bmi_name(c)
      character(kind=c_char), intent(in) :: c(*)
      integer :: i
      bmi_name = ' '
      do i = 1, 64
         if (c(i) == c_null_char) exit
         bmi_name(i:i) = c(i)
      enddo
   end function

   ! get_var_shape / get_var_type / get_var_rank (sfincs_bmi.f90:131-224)
   integer(c_int) function get_var_shape(c_var_name, var_shape) result(ierr) bind(C, name='get_var_shape')
      character(kind=c_char), intent(in) :: c_var_name(*)
      integer(c_int), intent(inout) :: var_shape(6)
      ierr = 0
      var_shape = 0
      select case (trim(bmi_name(c_var_name)))
      case ('z_xz', 'z_yz', 'zs', 'zsmax')
         var_shape(1) = host_np
      case ('zb')
         var_shape(1) = size(host_zb)
      case ('subgrid_z_zmin')
         var_shape(1) = size(host_zmin)
      case ('qext')
         if (allocated(host_qext)) var_shape(1) = size(host_qext)
      case ('z_index_z_n', 'z_index_z_m')
         if (allocated(host_zn)) var_shape(1) = size(host_zn)
      case default
         ierr = -1
      end select
   end function

   integer(c_int) function get_var_type(c_var_name, c_type) result(ierr) bind(C, name='get_var_type')
      character(kind=c_char), intent(in) :: c_var_name(*)
      character(kind=c_char), intent(out) :: c_type(64)
      character(8) :: tn
      integer :: i
      ierr = 0
      select case (trim(bmi_name(c_var_name)))
      case ('zs')
         tn = 'double'
      case ('z_xz', 'z_yz', 'zb', 'subgrid_z_zmin', 'zsmax', 'qext')
         tn = 'float'
      case ('z_index_z_n', 'z_index_z_m')
         tn = 'integer'
      case default
         tn = ' '
         ierr = -1
      end select
      c_type = c_null_char
      do i = 1, len_trim(tn)
         c_type(i) = tn(i:i)
      enddo
   end function

   integer(c_int) function get_var_rank(c_var_name, rank) result(ierr) bind(C, name='get_var_rank')
      character(kind=c_char), intent(in) :: c_var_name(*)
      integer(c_int), intent(out) :: rank
      ierr = 0
      rank = 1
      select case (trim(bmi_name(c_var_name)))
      case ('z_xz', 'z_yz', 'zs', 'zb', 'subgrid_z_zmin', 'zsmax', 'qext')
      case default
         ierr = -1
      end select
   end function

   ! get_sfincs_cell_index(es) and get_sfincs_cell_area (sfincs_bmi.f90:299-370)
   integer(c_int) function get_sfincs_cell_index(x, y, indx) result(ierr) bind(C, name='get_sfincs_cell_index')
      real(c_double), intent(in) :: x, y
      integer(c_int), intent(out) :: indx
      indx = host_find_cell(x, y)
      ierr = 0
   end function

   integer(c_int) function get_sfincs_cell_indices(x, y, indx, n) result(ierr) bind(C, name='get_sfincs_cell_indices')
      integer(c_int), value :: n
      real(c_double), intent(in) :: x(n), y(n)
      integer(c_int), intent(out) :: indx(n)
      integer :: i
      do i = 1, n
         indx(i) = host_find_cell(x(i), y(i))
      enddo
      ierr = 0
   end function

   integer(c_int) function get_sfincs_cell_area(indx, area) result(ierr) bind(C, name='get_sfincs_cell_area')
      integer(c_int), intent(in) :: indx
      real(c_double), intent(out) :: area
      ierr = 0
      area = 0.0d0
      if (indx < 1 .or. indx > size(host_iref)) then
         ierr = -1
      else
         area = real(host_area(host_iref(indx)), c_double)
      endif
   end function

   integer(c_int) function get_last_bmi_error(c_error) result(ierr) bind(C, name='get_last_bmi_error')
      character(kind=c_char), intent(out) :: c_error(1024)
      call host_last_error(c_error, 1024)
      ierr = 0
   end function

   ! set_logical switches qext (sfincs_bmi.f90:226-245). update_zbuv re-derives zbuvmx after
   ! a bed change (:277-284); wave-enhanced roughness, a process outside this build, returns
   ! -1 so that a coupler notices (:286-297).
   integer(c_int) function set_logical(c_flag_name, ival) result(ierr) bind(C, name='set_logical')
      character(kind=c_char), intent(in) :: c_flag_name(*)
      integer(c_int) :: ival(*)
      ierr = -1
      if (trim(bmi_name(c_flag_name)) == 'qext') ierr = host_qext_switch(ival(1) /= 0)
   end function

   integer(c_int) function update_zbuv() result(ierr) bind(C, name='update_zbuv')
      ierr = host_update_zbuv()
   end function

   integer(c_int) function update_apparent_roughness() result(ierr) bind(C, name='update_apparent_roughness')
      ierr = -1
   end function
end module sfincs_bmi
