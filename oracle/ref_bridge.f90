! ref_bridge.f90 -- C entry points onto the reference's own Fortran modules.
!
! TEST INFRASTRUCTURE ONLY.  Built by oracle/ref.mk together with the
! reference sources read in place (nothing of the reference is copied), into
! oracle/_ref/.  Each routine loads host arrays into a reference ParticleSet,
! calls the reference procedure unchanged and copies the result back, so the
! tests can hold the oracle against the reference itself.
module ref_bridge
  use iso_c_binding
  use iso_fortran_env, only: real64, int32, int8
  use mod_particles, only: ParticleSet
  use mod_particle_sorting, only: sort_particles_by_cell
  use mod_particleMover, only: move_particles_electrostatic
  use mod_particleBC, only: apply_particle_bc_legacy
  use mod_chargeDeposition, only: deposit_particle_set_to_np_thread
  use mod_charge_weights, only: build_kq
  use mod_electricField, only: calc_Efield_modular
  use mod_rng, only: ran2
  use mod_config, only: Config
  use mod_domain, only: Domain
  use mod_magneticField, only: MagneticField
  use mod_output_2d, only: write_plane_xy_scalar_2d, write_plane_xz_scalar_2d, &
                           write_plane_yz_scalar_2d
  implicit none
contains

  subroutine to_pset(p, np, x, y, z, vx, vy, vz, dead)
    type(ParticleSet), intent(inout) :: p
    integer(c_int64_t), intent(in) :: np
    real(c_double), intent(in) :: x(np), y(np), z(np), vx(np), vy(np), vz(np)
    integer(c_int8_t), intent(in) :: dead(np)
    call p%allocate_pset(int(max(np, 1_c_int64_t), int32))
    p%n = int(np, int32)
    if (np > 0) then
      p%x(1:np) = x; p%y(1:np) = y; p%z(1:np) = z
      p%vx(1:np) = vx; p%vy(1:np) = vy; p%vz(1:np) = vz
      p%flag_dead(1:np) = int(dead, int8)
    end if
  end subroutine to_pset

  subroutine from_pset(p, np, x, y, z, vx, vy, vz, dead)
    type(ParticleSet), intent(in) :: p
    integer(c_int64_t), intent(inout) :: np
    real(c_double), intent(inout) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
    integer(c_int8_t), intent(inout) :: dead(*)
    integer :: m
    m = p%n
    np = m
    if (m > 0) then
      x(1:m) = p%x(1:m); y(1:m) = p%y(1:m); z(1:m) = p%z(1:m)
      vx(1:m) = p%vx(1:m); vy(1:m) = p%vy(1:m); vz(1:m) = p%vz(1:m)
      dead(1:m) = int(p%flag_dead(1:m), c_int8_t)
    end if
  end subroutine from_pset

  real(c_double) function ref_ran2(irand) bind(C, name="ref_ran2")
    integer(c_int32_t), intent(inout) :: irand
    ref_ran2 = ran2(irand)
  end function ref_ran2

  subroutine ref_sort(n, h, np, x, y, z, vx, vy, vz, dead, cell, ccount, cstart) bind(C, name="ref_sort")
    integer(c_int32_t), intent(in) :: n(3)
    real(c_double), intent(in) :: h(3)
    integer(c_int64_t), value :: np
    real(c_double), intent(inout) :: x(np), y(np), z(np), vx(np), vy(np), vz(np)
    integer(c_int8_t), intent(inout) :: dead(np)
    integer(c_int32_t), intent(out) :: cell(np), ccount(*), cstart(*)
    type(ParticleSet) :: p
    integer(c_int64_t) :: m
    integer :: nc
    call to_pset(p, np, x, y, z, vx, vy, vz, dead)
    call sort_particles_by_cell(p, n, h)
    m = np
    call from_pset(p, m, x, y, z, vx, vy, vz, dead)
    nc = n(1) * n(2) * n(3)
    if (np > 0) cell(1:np) = p%cell_id(1:np)
    ccount(1:nc) = p%cell_count(1:nc)
    cstart(1:nc) = p%cell_start(1:nc)
  end subroutine ref_sort

  subroutine ref_move_es(n, h, np, x, y, z, vx, vy, vz, E3, q, m, dt) bind(C, name="ref_move_es")
    integer(c_int32_t), intent(in) :: n(3)
    real(c_double), intent(in) :: h(3)
    integer(c_int64_t), value :: np
    real(c_double), intent(inout) :: x(np), y(np), z(np), vx(np), vy(np), vz(np)
    real(c_double), intent(in) :: E3(3, 0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    real(c_double), value :: q, m, dt
    type(ParticleSet) :: p
    integer(c_int8_t), allocatable :: dead(:)
    integer(c_int64_t) :: k
    allocate(dead(max(np, 1_c_int64_t)))
    dead = 0
    call to_pset(p, np, x, y, z, vx, vy, vz, dead)
    call move_particles_electrostatic(p, n, h, E3, q, m, dt)
    k = np
    call from_pset(p, k, x, y, z, vx, vy, vz, dead)
  end subroutine ref_move_es

  subroutine ref_apply_bc(n, h, np, x, y, z, vx, vy, vz, dead, bcnd, flag_pbc, flag_nmn, ptype, &
                          tag_neg, qmacro, mass, Nm, ngrid, p_mac, P_loss) bind(C, name="ref_apply_bc")
    integer(c_int32_t), intent(in) :: n(3)
    real(c_double), intent(in) :: h(3)
    integer(c_int64_t), intent(inout) :: np
    real(c_double), intent(inout) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
    integer(c_int8_t), intent(inout) :: dead(*)
    integer(c_int32_t), intent(in) :: bcnd(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    integer(c_int32_t), value :: flag_pbc, flag_nmn, ptype, tag_neg, ngrid
    real(c_double), value :: qmacro, mass, Nm
    real(c_double), intent(inout) :: p_mac(2, ngrid+1)
    real(c_double), intent(inout) :: P_loss
    type(ParticleSet) :: p
    integer(int32) :: dtype(ngrid+1)
    real(real64), allocatable :: sxz(:,:), syz(:,:,:)
    allocate(sxz(0:n(1)+2, 0:n(3)+2), syz(2, 0:n(2)+2, 0:n(3)+2))
    sxz = 0.0_real64
    syz = 0.0_real64
    dtype = 0
    call to_pset(p, np, x(1:np), y(1:np), z(1:np), vx(1:np), vy(1:np), vz(1:np), dead(1:np))
    call apply_particle_bc_legacy(p, n, h, bcnd, real(n(1),real64)*h(1), real(n(2),real64)*h(2), &
         real(n(3),real64)*h(3), flag_pbc, flag_nmn, ptype, tag_neg, 0_int32, dtype, qmacro, &
         sxz, syz, p_mac, mass, P_loss, Nm)
    call from_pset(p, np, x, y, z, vx, vy, vz, dead)
  end subroutine ref_apply_bc

  ! apply_particle_bc_legacy with dielectric charging: dtype_in over labels
  ! 0..ngrid handed over as State hands dom%dtype(0:ngrid) to the routine.
  subroutine ref_apply_bc_die(n, h, np, x, y, z, vx, vy, vz, dead, bcnd, flag_pbc, flag_nmn, &
                              ptype, tag_neg, qmacro, mass, Nm, ngrid, dtype_in, sxz, syz, &
                              p_mac, P_loss) bind(C, name="ref_apply_bc_die")
    integer(c_int32_t), intent(in) :: n(3)
    real(c_double), intent(in) :: h(3)
    integer(c_int64_t), intent(inout) :: np
    real(c_double), intent(inout) :: x(*), y(*), z(*), vx(*), vy(*), vz(*)
    integer(c_int8_t), intent(inout) :: dead(*)
    integer(c_int32_t), intent(in) :: bcnd(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    integer(c_int32_t), value :: flag_pbc, flag_nmn, ptype, tag_neg, ngrid
    real(c_double), value :: qmacro, mass, Nm
    integer(c_int32_t), intent(in) :: dtype_in(0:ngrid)
    real(c_double), intent(inout) :: sxz(0:n(1)+2, 0:n(3)+2), syz(2, 0:n(2)+2, 0:n(3)+2)
    real(c_double), intent(inout) :: p_mac(2, ngrid+1)
    real(c_double), intent(inout) :: P_loss
    type(ParticleSet) :: p
    integer(int32), allocatable :: dt(:)
    allocate(dt(0:ngrid))
    dt = dtype_in
    call to_pset(p, np, x(1:np), y(1:np), z(1:np), vx(1:np), vy(1:np), vz(1:np), dead(1:np))
    call apply_particle_bc_legacy(p, n, h, bcnd, real(n(1),real64)*h(1), real(n(2),real64)*h(2), &
         real(n(3),real64)*h(3), flag_pbc, flag_nmn, ptype, tag_neg, 1_int32, dt, qmacro, &
         sxz, syz, p_mac, mass, P_loss, Nm)
    call from_pset(p, np, x, y, z, vx, vy, vz, dead)
  end subroutine ref_apply_bc_die

  subroutine ref_deposit(n, h, np, x, y, z, dead, kq, Nm, np_local) bind(C, name="ref_deposit")
    integer(c_int32_t), intent(in) :: n(3)
    real(c_double), intent(in) :: h(3)
    integer(c_int64_t), value :: np
    real(c_double), intent(in) :: x(np), y(np), z(np)
    integer(c_int8_t), intent(in) :: dead(np)
    real(c_double), intent(in) :: kq(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    real(c_double), value :: Nm
    real(c_double), intent(inout) :: np_local(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    type(ParticleSet) :: p
    real(c_double), allocatable :: zero(:)
    allocate(zero(max(np, 1_c_int64_t)))
    zero = 0.0_c_double
    call to_pset(p, np, x, y, z, zero, zero, zero, dead)
    call deposit_particle_set_to_np_thread(p, n, h, kq, Nm, np_local)
  end subroutine ref_deposit

  subroutine ref_build_kq(n, bcnd, kq) bind(C, name="ref_build_kq")
    integer(c_int32_t), intent(in) :: n(3)
    integer(c_int32_t), intent(in) :: bcnd(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    real(c_double), intent(out) :: kq(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    integer, allocatable :: b(:,:,:)
    allocate(b(0:n(1)+2, 0:n(2)+2, 0:n(3)+2))
    b = bcnd
    call build_kq(b, kq)
  end subroutine ref_build_kq

  subroutine ref_efield(n, h, phi, E3, bcnd) bind(C, name="ref_efield")
    integer(c_int32_t), intent(in) :: n(3)
    real(c_double), intent(in) :: h(3)
    real(c_double), intent(in) :: phi(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    real(c_double), intent(inout) :: E3(3, 0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    integer(c_int32_t), intent(in) :: bcnd(0:n(1)+2, 0:n(2)+2, 0:n(3)+2)
    call calc_Efield_modular(n, h, phi, E3, bcnd)
  end subroutine ref_efield

  ! MagneticField%build_from_cfg on a Config holding nB field records
  ! (file/name/info as blank-padded characters) and a Domain holding the box
  ! extents in m; Bi comes back as (4, 0:nxB+2, 0:nyB+2, 0:nzB+2).
  subroutine ref_build_bfield(nB, n_B, files, names, infos, scale, B0, dL, x0, y0, z0, box, &
                              Bi, flag_thr) bind(C, name="ref_build_bfield")
    integer(c_int32_t), value :: nB
    integer(c_int32_t), intent(in) :: n_B(3)
    character(kind=c_char), intent(in) :: files(5), names(20, 5), infos(2, 5)
    real(c_double), intent(in) :: scale(5), B0(5), dL(5), x0(5), y0(5), z0(5), box(3)
    real(c_double), intent(out) :: Bi(*)
    integer(c_int32_t), intent(out) :: flag_thr
    type(Config) :: cfg
    type(Domain) :: dom
    type(MagneticField) :: mf
    integer :: i, k, m
    cfg%nB = nB
    cfg%n_B = n_B
    cfg%flag_B = 0
    do i = 1, 5
      cfg%B_file(i) = files(i)
      do k = 1, 20
        cfg%B_name(i)(k:k) = names(k, i)
      end do
      do k = 1, 2
        cfg%B_info(i)(k:k) = infos(k, i)
      end do
      if (i <= nB .and. cfg%B_file(i) /= 's' .and. cfg%B_file(i) /= 'S') cfg%flag_B = 1
    end do
    cfg%B_scale = scale; cfg%B0 = B0; cfg%dL = dL
    cfg%x0 = x0; cfg%y0 = y0; cfg%z0 = z0
    dom%xmax = box(1); dom%ymax = box(2); dom%zmax = box(3)
    call mf%build_from_cfg(cfg, dom, 1)
    m = size(mf%Bi)
    Bi(1:m) = reshape(mf%Bi, [m])
    flag_thr = mf%flag_thr
    call mf%destroy()
  end subroutine ref_build_bfield

  ! mod_output_2d's plane writers: kind 1 xy (f over x, y), 2 xz, 3 yz.
  subroutine ref_write_plane_2d(kind, fname, nlen, f, n, every) bind(C, name="ref_write_plane_2d")
    integer(c_int32_t), value :: kind, nlen, every
    character(kind=c_char), intent(in) :: fname(nlen)
    real(c_double), intent(in) :: f(*)
    integer(c_int32_t), intent(in) :: n(3)
    character(len=nlen) :: name
    integer :: i
    do i = 1, nlen
      name(i:i) = fname(i)
    end do
    select case (kind)
    case (1)
      call write_plane_xy_scalar_2d(name, f, n, every)
    case (2)
      call write_plane_xz_scalar_2d(name, f, n, every)
    case default
      call write_plane_yz_scalar_2d(name, f, n, every)
    end select
  end subroutine ref_write_plane_2d

  ! An ESw.d or Fw.d edit of v into buf (kind 1 ES, 2 F), blank-padded.
  subroutine ref_format(kind, v, w, d, buf, blen) bind(C, name="ref_format")
    integer(c_int32_t), value :: kind, w, d, blen
    real(c_double), value :: v
    character(kind=c_char), intent(out) :: buf(blen)
    character(len=32) :: fmt
    character(len=64) :: s
    integer :: i
    if (kind == 1) then
      write(fmt, '(a,i0,a,i0,a)') '(ES', w, '.', d, ')'
    else
      write(fmt, '(a,i0,a,i0,a)') '(F', w, '.', d, ')'
    end if
    s = ' '
    write(s, fmt) v
    do i = 1, blen
      buf(i) = s(i:i)
    end do
  end subroutine ref_format

  ! The particle backup as Src/main.f90:1321-1327 writes it for one thread
  ! (kind 1) or reads it as Src/restart.f90:63-70 (kind 2): the time, np_tot,
  ! each species' (6, np) block, then flag_cex of species tag_neg if > 0.
  subroutine ref_backup_io(kind, fname, nlen, time, ntype, np_tot, vxp, tag_neg, cex) &
      bind(C, name="ref_backup_io")
    integer(c_int32_t), value :: kind, nlen, ntype, tag_neg
    character(kind=c_char), intent(in) :: fname(nlen)
    real(c_double), intent(inout) :: time
    integer(c_int32_t), intent(inout) :: np_tot(ntype)
    real(c_double), intent(inout) :: vxp(6, *)
    integer(c_int32_t), intent(inout) :: cex(*)
    character(len=nlen) :: name
    integer :: i, u, t, off
    do i = 1, nlen
      name(i:i) = fname(i)
    end do
    if (kind == 1) then
      open(newunit=u, file=name, form='UNFORMATTED')
      write(u) time
      write(u) np_tot(1:ntype)
      off = 0
      do t = 1, ntype
        write(u) vxp(:, off + 1:off + np_tot(t))
        off = off + np_tot(t)
      end do
      if (tag_neg > 0) write(u) cex(1:np_tot(tag_neg))
      close(u)
    else
      open(newunit=u, file=name, form='UNFORMATTED', status='old')
      read(u) time
      read(u) np_tot(1:ntype)
      off = 0
      do t = 1, ntype
        read(u) vxp(:, off + 1:off + np_tot(t))
        off = off + np_tot(t)
      end do
      if (tag_neg > 0) read(u) cex(1:np_tot(tag_neg))
      close(u)
    end if
  end subroutine ref_backup_io

end module ref_bridge
