! ORACLE -- test infrastructure only (see oracle/sfx_oracle.c header).
!
! Thin C-callable driver around the reference's own hot-path modules, compiled from
! /root/reference/source/src by oracle/Makefile into oracle/_ref/libsfincs_ref.so:
!   sfincs_data        (module state the kernels read and write)
!   sfincs_momentum    compute_fluxes            (sfincs_momentum.f90:9-784)
!   sfincs_continuity  compute_water_levels      (sfincs_continuity.f90:5-43)
!   sfincs_input       read_sfincs_input         (sfincs_input.f90:5-747, defaults)
! The driver fills the module arrays from caller buffers, runs one routine and copies
! results back. It contains no physics of its own.
module ref_driver
   use iso_c_binding
   use sfincs_data
   use sfincs_momentum
   use sfincs_continuity
   use sfincs_timestep_analysis
   use sfincs_input
   use sfincs_log
   use sfincs_spiderweb
   implicit none
   real :: tloopflux = 0.0
   real :: tloopcont = 0.0
contains

   function cstr(s) result(f)
      character(kind=c_char), intent(in) :: s(*)
      character(len=64) :: f
      integer :: i
      f = ' '
      do i = 1, 64
         if (s(i) == c_null_char) exit
         f(i:i) = s(i)
      enddo
   end function

   subroutine ref_set_sizes(np_, npuv_, ncuv_, nref_, nl_) bind(C, name='ref_set_sizes')
      integer(c_int32_t), value :: np_, npuv_, ncuv_, nref_, nl_
      np = np_; npuv = npuv_; ncuv = ncuv_; nref = int(nref_, 1); subgrid_nlevels = nl_
      ! switches the hot path reads that this harness never turns on
      crsgeo = .false.; patmos = .false.; snapwave = .false.; timestep_analysis = .false.
      wave_enhanced_roughness = .false.; use_qext = .false.; wavemaker = .false.
      store_t_zsmax = .false.; store_maximum_velocity = .false.; store_maximum_flux = .false.
      store_twet = .false.; use_storage_volume = .false.; h73table = .false.
      nsrcdrn = 0
   end subroutine

   subroutine ref_set_scalar(name, v) bind(C, name='ref_set_scalar')
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), value :: v
      select case (trim(cstr(name)))
      case ('subgrid');            subgrid = (v /= 0.0d0)
      case ('use_quadtree');       use_quadtree = (v /= 0.0d0)
      case ('advection');          advection = (v /= 0.0d0)
      case ('advection_scheme');   advection_scheme = int(v)
      case ('viscosity');          viscosity = (v /= 0.0d0)
      case ('coriolis');           coriolis = (v /= 0.0d0)
      case ('friction2d');         friction2d = (v /= 0.0d0)
      case ('thetasmoothing');     thetasmoothing = (v /= 0.0d0)
      case ('wiggle_suppression'); wiggle_suppression = (v /= 0.0d0)
      case ('store_zsmax');        store_maximum_waterlevel = (v /= 0.0d0)
      case ('store_velmax');       store_maximum_velocity = (v /= 0.0d0)
      case ('store_fluxmax');      store_maximum_flux = (v /= 0.0d0)
      case ('store_twet');         store_twet = (v /= 0.0d0)
      case ('twet_threshold');     twet_threshold = real(v, 4)
      case ('use_qext');           use_qext = (v /= 0.0d0)
      case ('use_storage_volume'); use_storage_volume = (v /= 0.0d0)
      case ('timestep_analysis')   ! after dtmax and the sizes (initialize_timestep_analysis)
         timestep_analysis = (v /= 0.0d0)
         if (timestep_analysis) then
            if (allocated(timestep_analysis_average_required_timestep)) then
               deallocate(timestep_analysis_average_required_timestep, timestep_analysis_required_timestep, &
                          timestep_analysis_times_limiting, timestep_analysis_times_wet)
            endif
            call initialize_timestep_analysis()
         endif
      case ('wind');               wind = (v /= 0.0d0)
      case ('precip');             precip = (v /= 0.0d0)
      case ('patmos');             patmos = (v /= 0.0d0)
      case ('h73table');           h73table = (v /= 0.0d0)
      case ('nsrc');               nsrcdrn = int(v)
      case ('g');                  g = real(v, 4)
      case ('huthresh');           huthresh = real(v, 4)
      case ('huvmin');             huvmin = real(v, 4)
      case ('advlim');             advlim = real(v, 4)
      case ('slopelim');           slopelim = real(v, 4)
      case ('uvlim');              uvlim = real(v, 4)
      case ('nuviscfac');          nuviscfac = real(v, 4)
      case ('fcorio');             fcorio = real(v, 4)
      case ('theta');              theta = real(v, 4)
      case ('wiggle_factor');      wiggle_factor = real(v, 4)
      case ('wiggle_threshold');   wiggle_threshold = real(v, 4)
      case ('dtmax');              dtmax = real(v, 4)
      case ('alpha');              alfa = real(v, 4)
      case ('rhow');               rhow = real(v, 4)
      case ('dx');                 dx = real(v, 4)
      case ('dy');                 dy = real(v, 4)
      case default
      end select
   end subroutine

   subroutine ref_set_i32(name, p, n) bind(C, name='ref_set_i32')
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int32_t), value :: n
      integer(c_int32_t), intent(in) :: p(n)
      select case (trim(cstr(name)))
      case ('uv_index_z_nm');   call put4(uv_index_z_nm)
      case ('uv_index_z_nmu');  call put4(uv_index_z_nmu)
      case ('uv_index_u_nmd');  call put4(uv_index_u_nmd)
      case ('uv_index_u_nmu');  call put4(uv_index_u_nmu)
      case ('uv_index_u_ndm');  call put4(uv_index_u_ndm)
      case ('uv_index_u_num');  call put4(uv_index_u_num)
      case ('uv_index_v_ndm');  call put4(uv_index_v_ndm)
      case ('uv_index_v_nm');   call put4(uv_index_v_nm)
      case ('uv_index_v_ndmu'); call put4(uv_index_v_ndmu)
      case ('uv_index_v_nmu');  call put4(uv_index_v_nmu)
      case ('cuv_index_uv');    call put4(cuv_index_uv)
      case ('cuv_index_uv1');   call put4(cuv_index_uv1)
      case ('cuv_index_uv2');   call put4(cuv_index_uv2)
      case ('z_index_uv_md');   call put4(z_index_uv_md)
      case ('z_index_uv_mu');   call put4(z_index_uv_mu)
      case ('z_index_uv_nd');   call put4(z_index_uv_nd)
      case ('z_index_uv_nu');   call put4(z_index_uv_nu)
      case ('z_index_uv_md1');  call put4(z_index_uv_md1)
      case ('z_index_uv_md2');  call put4(z_index_uv_md2)
      case ('z_index_uv_mu1');  call put4(z_index_uv_mu1)
      case ('z_index_uv_mu2');  call put4(z_index_uv_mu2)
      case ('z_index_uv_nd1');  call put4(z_index_uv_nd1)
      case ('z_index_uv_nd2');  call put4(z_index_uv_nd2)
      case ('z_index_uv_nu1');  call put4(z_index_uv_nu1)
      case ('z_index_uv_nu2');  call put4(z_index_uv_nu2)
      case ('nmindsrc');        call put4(nmindsrc)
      case default
      end select
   contains
      subroutine put4(a)
         integer*4, allocatable :: a(:)
         if (allocated(a)) deallocate(a)
         allocate(a(n))
         a = p
      end subroutine
   end subroutine

   subroutine ref_set_i8(name, p, n) bind(C, name='ref_set_i8')
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int32_t), value :: n
      integer(c_int8_t), intent(in) :: p(n)
      select case (trim(cstr(name)))
      case ('kcs');           call put1(kcs)
      case ('kcuv');          call put1(kcuv)
      case ('kfuv');          call put1(kfuv)
      case ('mask_adv');      call put1(mask_adv)
      case ('z_flags_iref');  call put1(z_flags_iref)
      case ('uv_flags_iref'); call put1(uv_flags_iref)
      case ('uv_flags_type'); call put1(uv_flags_type)
      case ('uv_flags_dir');  call put1(uv_flags_dir)
      case default
      end select
   contains
      subroutine put1(a)
         integer*1, allocatable :: a(:)
         if (allocated(a)) deallocate(a)
         allocate(a(n))
         a = p
      end subroutine
   end subroutine

   subroutine ref_set_f32(name, p, n) bind(C, name='ref_set_f32')
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int32_t), value :: n
      real(c_float), intent(in) :: p(n)
      select case (trim(cstr(name)))
      case ('zb');           call putr(zb)
      case ('zbuvmx');       call putr(zbuvmx)
      case ('gn2uv');        call putr(gn2uv)
      case ('dxrinv');       call putr(dxrinv)
      case ('dyrinv');       call putr(dyrinv)
      case ('dxr2inv');      call putr(dxr2inv)
      case ('dyr2inv');      call putr(dyr2inv)
      case ('dxrinvc');      call putr(dxrinvc)
      case ('dyrinvc');      call putr(dyrinvc)
      case ('dxrm');         call putr(dxrm)
      case ('dyrm');         call putr(dyrm)
      case ('cell_area');    call putr(cell_area)
      case ('nuvisc');       call putr(nuvisc)
      case ('q');            call putr(q)
      case ('q0');           call putr(q0)
      case ('uv0');          call putr(uv0)
      case ('uv');           call putr(uv)
      case ('zs0');          call putr(zs0)
      case ('zsderv');       call putr(zsderv)
      case ('zsmax');        call putr(zsmax)
      case ('vmax');         call putr(vmax)
      case ('qmax');         call putr(qmax)
      case ('twet');         call putr(twet)
      case ('qext');         call putr(qext)
      case ('storage_volume'); call putr(storage_volume)
      case ('tauwu');        call putr(tauwu)
      case ('tauwv');        call putr(tauwv)
      case ('netprcp');      call putr(netprcp)
      case ('patm');         call putr(patm)
      case ('qtsrc');        call putr(qtsrc)
      case ('z_zmin');       call putr(subgrid_z_zmin)
      case ('z_zmax');       call putr(subgrid_z_zmax)
      case ('z_volmax');     call putr(subgrid_z_volmax)
      case ('uv_zmin');      call putr(subgrid_uv_zmin)
      case ('uv_zmax');      call putr(subgrid_uv_zmax)
      case ('uv_navg_w');    call putr(subgrid_uv_navg_w)
      case ('uv_fnfit');     call putr(subgrid_uv_fnfit)
      case ('uv_havg_zmax'); call putr(subgrid_uv_havg_zmax)
      case ('uv_nrep_zmax'); call putr(subgrid_uv_nrep_zmax)
      case ('z_dep');        call putr2(subgrid_z_dep)
      case ('uv_havg');      call putr2(subgrid_uv_havg)
      case ('uv_nrep');      call putr2(subgrid_uv_nrep)
      case ('uv_pwet');      call putr2(subgrid_uv_pwet)
      case default
      end select
   contains
      subroutine putr(a)
         real*4, allocatable :: a(:)
         if (allocated(a)) then
            if (size(a) /= n) deallocate(a)
         endif
         if (.not. allocated(a)) allocate(a(n))
         a = p
      end subroutine
      subroutine putr2(a)
         real*4, allocatable :: a(:,:)
         integer :: nlv
         nlv = subgrid_nlevels
         if (allocated(a)) deallocate(a)
         allocate(a(nlv, n / nlv))
         a = reshape(p, (/ nlv, n / nlv /))
      end subroutine
   end subroutine

   subroutine ref_set_f64(name, p, n) bind(C, name='ref_set_f64')
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: p(n)
      select case (trim(cstr(name)))
      case ('zs');       call putd(zs)
      case ('z_volume'); call putd(z_volume)
      case default
      end select
   contains
      subroutine putd(a)
         real*8, allocatable :: a(:)
         if (allocated(a)) then
            if (size(a) /= n) deallocate(a)
         endif
         if (.not. allocated(a)) allocate(a(n))
         a = p
      end subroutine
   end subroutine

   subroutine ref_get(name, p, nbytes) bind(C, name='ref_get')
      character(kind=c_char), intent(in) :: name(*)
      type(c_ptr), value :: p
      integer(c_int64_t), value :: nbytes
      real(c_float), pointer :: f(:)
      real(c_double), pointer :: d(:)
      integer(c_int8_t), pointer :: b(:)
      integer(c_int32_t), pointer :: iv(:)
      select case (trim(cstr(name)))
      case ('q');        call c_f_pointer(p, f, [nbytes / 4]); f = q
      case ('uv');       call c_f_pointer(p, f, [nbytes / 4]); f = uv
      case ('zs0');      call c_f_pointer(p, f, [nbytes / 4]); f = zs0
      case ('zsderv');   call c_f_pointer(p, f, [nbytes / 4]); f = zsderv
      case ('zsmax');    call c_f_pointer(p, f, [nbytes / 4]); f = zsmax
      case ('vmax');     call c_f_pointer(p, f, [nbytes / 4]); f = vmax
      case ('qmax');     call c_f_pointer(p, f, [nbytes / 4]); f = qmax
      case ('twet');     call c_f_pointer(p, f, [nbytes / 4]); f = twet
      case ('storage_volume'); call c_f_pointer(p, f, [nbytes / 4]); f = storage_volume
      case ('tsa_sum');  call c_f_pointer(p, f, [nbytes / 4]); f = timestep_analysis_average_required_timestep
      case ('tsa_wet');  call c_f_pointer(p, iv, [nbytes / 4]); iv = timestep_analysis_times_wet
      case ('tsa_lim');  call c_f_pointer(p, iv, [nbytes / 4]); iv = timestep_analysis_times_limiting
      case ('zs');       call c_f_pointer(p, d, [nbytes / 8]); d = zs
      case ('z_volume'); call c_f_pointer(p, d, [nbytes / 8]); d = z_volume
      case ('kfuv');     call c_f_pointer(p, b, [nbytes]);     b = kfuv
      case ('x73');      call c_f_pointer(p, f, [nbytes / 4]); f = reshape(x73, [10000])
      case default
      end select
   end subroutine

   ! timestep_analysis_finalize(nt) (sfincs_timestep_analysis.f90:77): the per-cell fields
   subroutine ref_tsa_finalize(nt_, dtc, pct) bind(C, name='ref_tsa_finalize')
      integer(c_int32_t), value :: nt_
      real(c_float), intent(out) :: dtc(np), pct(np)
      call timestep_analysis_finalize(int(nt_))
      dtc = timestep_analysis_average_required_timestep_per_cell
      pct = timestep_analysis_percentage_limiting_per_cell
      deallocate(timestep_analysis_average_required_timestep_per_cell, timestep_analysis_percentage_limiting_per_cell)
   end subroutine

   ! compute_fluxes(dt, tloop) (sfincs_momentum.f90:9); returns min_dt
   function ref_compute_fluxes(dt_) result(mdt) bind(C, name='ref_compute_fluxes')
      real(c_float), value :: dt_
      real(c_float) :: mdt
      real*4 :: dtl
      dtl = dt_
      call compute_fluxes(dtl, tloopflux)
      if (timestep_analysis) call timestep_analysis_update(min_dt)   ! sfincs_lib.f90:588
      mdt = min_dt
   end function

   ! compute_water_levels(t, dt, tloop) (sfincs_continuity.f90:5)
   subroutine ref_compute_water_levels(t_, dt_) bind(C, name='ref_compute_water_levels')
      real(c_double), value :: t_
      real(c_float), value :: dt_
      real*8 :: tl
      real*4 :: dtl
      tl = t_; dtl = dt_
      call compute_water_levels(tl, dtl, tloopcont)
   end subroutine

   ! the reference's own per-phase timers (sfincs_lib.f90:743,756)
   subroutine ref_timers(tf, tc) bind(C, name='ref_timers')
      real(c_double), intent(out) :: tf, tc
      tf = tloopflux; tc = tloopcont
   end subroutine

   subroutine ref_reset_timers() bind(C, name='ref_reset_timers')
      tloopflux = 0.0; tloopcont = 0.0
   end subroutine

   ! fill_h73_tables (sfincs_data.f90:858-881)
   subroutine ref_fill_h73() bind(C, name='ref_fill_h73')
      h73table = .true.
      if (allocated(x73)) deallocate(x73)
      call fill_h73_tables()
   end subroutine

   ! The reference's own ASCII meteo readers (sfincs_spiderweb.f90): dims(1:4) = nt, nrows,
   ! ncols, nquant and geo(1:5) = x_llcorner, y_llcorner, dx, dy, spw_radius; when cap holds
   ! them, times(nt) and the fields, [it][row][col] with the row index slowest after time
   ! (spiderweb: vmag, vdir, pdrp, prcp one after the other).
   subroutine ref_read_meteo(kind, path, tref, dims, geo, times, vals, cap) bind(C, name='ref_read_meteo')
      integer(c_int32_t), value :: kind, cap
      character(kind=c_char), intent(in) :: path(256), tref(15)
      integer(c_int32_t), intent(out) :: dims(4)
      real(c_float), intent(out) :: geo(5)
      real(c_float), intent(out) :: times(*), vals(*)
      character(len=256) :: fname
      character(len=15) :: trefstr
      integer :: nt, nrows, ncols, nquant, i, it, n, m, k
      real*4 :: xll, yll, dx, dy, rad
      real*4, allocatable :: a(:,:,:), b(:,:,:), c(:,:,:), d(:,:,:), tt(:), xe(:), ye(:)
      fname = ' '
      do i = 1, 256
         if (path(i) == c_null_char) exit
         fname(i:i) = path(i)
      enddo
      do i = 1, 15
         trefstr(i:i) = tref(i)
      enddo
      xll = 0.0; yll = 0.0; dx = 0.0; dy = 0.0; rad = 0.0
      if (kind == 0) then
         call read_amuv_dimensions(fname, nt, nrows, ncols, xll, yll, dx, dy, nquant)
      else
         call read_spw_dimensions(fname, nt, nrows, ncols, rad, nquant)
      endif
      dims = [nt, nrows, ncols, nquant]
      geo = [xll, yll, dx, dy, rad]
      if (cap < nt * nrows * ncols * merge(1, 4, kind == 0) + merge(0, nt, kind == 0)) return
      allocate(a(nt, nrows, ncols), tt(nt))
      if (kind == 0) then
         call read_amuv_file(fname, nt, nrows, ncols, tt, a, trefstr)
      else
         allocate(b(nt, nrows, ncols), c(nt, nrows, ncols), d(nt, nrows, ncols), xe(nt), ye(nt))
         call read_spw_file(fname, nt, nrows, ncols, rad, tt, xe, ye, a, b, c, d, nquant, trefstr)
      endif
      times(1:nt) = tt
      k = 0
      do it = 1, nt
         do n = 1, nrows
            do m = 1, ncols
               k = k + 1
               vals(k) = a(it, n, m)
            enddo
         enddo
      enddo
      if (kind /= 0) then
         do it = 1, nt
            do n = 1, nrows
               do m = 1, ncols
                  vals(k + 1) = b(it, n, m)
                  vals(k + 1 + nt * nrows * ncols) = c(it, n, m)
                  vals(k + 1 + 2 * nt * nrows * ncols) = d(it, n, m)
                  k = k + 1
               enddo
            enddo
         enddo
         times(nt + 1:2 * nt) = xe
         ! the eye's y after the four fields (times holds the eye's x after the times)
         vals(4 * nt * nrows * ncols + 1:4 * nt * nrows * ncols + nt) = ye
      endif
   end subroutine

   ! read_sfincs_input (sfincs_input.f90:5) from ./sfincs.inp; returns the values the hot
   ! path depends on, in the order of ref_input_names.
   subroutine ref_read_input(out, nout) bind(C, name='ref_read_input')
      integer(c_int32_t), value :: nout
      real(c_double), intent(out) :: out(nout)
      call open_log()
      call read_sfincs_input()
      call close_log()
      out = 0.0d0
      out(1) = alfa; out(2) = theta; out(3) = hmin_cfl; out(4) = manning; out(5) = manning_land
      out(6) = manning_sea; out(7) = rghlevland; out(8) = zini; out(9) = dtmax; out(10) = huthresh
      out(11) = huvmin; out(12) = rhoa; out(13) = rhow; out(14) = advlim; out(15) = slopelim
      out(16) = btfilter; out(17) = btrelax; out(18) = wiggle_factor; out(19) = wiggle_threshold
      out(20) = uvlim; out(21) = uvmax; out(22) = nuviscdim; out(23) = nuviscfac; out(24) = latitude
      out(25) = g; out(26) = pi
      out(27) = merge(1.0d0, 0.0d0, advection); out(28) = advection_scheme
      out(29) = merge(1.0d0, 0.0d0, viscosity); out(30) = merge(1.0d0, 0.0d0, coriolis)
      out(31) = merge(1.0d0, 0.0d0, friction2d); out(32) = merge(1.0d0, 0.0d0, advection_mask)
      out(33) = merge(1.0d0, 0.0d0, wiggle_suppression); out(34) = merge(1.0d0, 0.0d0, store_maximum_waterlevel)
      out(35) = bndtype; out(36) = merge(1.0d0, 0.0d0, h73table)
      out(37) = t0; out(38) = t1; out(39) = tspinup; out(40) = fcorio
      out(41) = merge(1.0d0, 0.0d0, thetasmoothing)
      out(42) = qinf; out(43) = qinf_zmin
   end subroutine

end module ref_driver
