!> iso_c_binding interface of the MI355X non-bonded energy/force path (include/molsim_gpu.h).
!!
!! A MOLSIM build uses this module at the three places its non-bonded path is reached
!! (UTotal, Src/energy.F90:280; DUTotal, Src/denergy.F90:98; MCUpdate, Src/mc.F90:5203) and
!! passes its own arrays straight through: ro(3,np), az(na), iptpn(np), ubuf(nbuf),
!! iubuflow(natat) (1-based, as MOLSIM holds them), force(3,na).  Nonzero status codes are
!! turned into a stop with the library's message by GpuCheck, as MOLSIM's Stop does.
module MolsimGpu

   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: MSG_OK = 0
   integer(c_int), parameter :: MSG_LIST_ALL = 0, MSG_LIST_CELL = 1, MSG_LIST_VERLET = 2
   integer(c_int), parameter :: MSG_EWALD_NONE = 0, MSG_EWALD_STD = 1, MSG_EWALD_SPM = 2
   integer(c_int), parameter :: MSG_PREC_MIXED = 0, MSG_PREC_FP64 = 1

   type, bind(C) :: msg_params
      integer(c_int) :: np, npt
      real(c_double) :: boxlen(3)
      real(c_double) :: rcut, drnlist, epsi4pi
      integer(c_int) :: list_mode, ewald_mode, lsurf
      real(c_double) :: alpha
      integer(c_int) :: ncut, ncutregion, nmesh, order, precision
      real(c_double) :: force_scale
   end type msg_params

   type, bind(C) :: msg_pot
      integer(c_int) :: npt, form, lewald, pad
      real(c_double) :: alpha, rcut, epsi4pi
      real(c_double) :: utoltab, ftoltab, umaxtab, fmaxtab, r2uminin
      real(c_double) :: z(8), sig(8), eps(8), rad(8)
   end type msg_pot

   interface

      function msg_last_error() bind(C, name='msg_last_error')
         import :: c_ptr
         type(c_ptr) :: msg_last_error
      end function

      integer(c_int) function msg_build_table(pot, ubuf, maxbuf, iubuflow, nugrid, r2umin, r2atat) &
            bind(C, name='msg_build_table')
         import :: c_int, c_double, msg_pot
         type(msg_pot), intent(in) :: pot
         real(c_double), intent(out) :: ubuf(*)
         integer(c_int), value :: maxbuf
         integer(c_int), intent(out) :: iubuflow(*), nugrid(*)
         real(c_double), intent(out) :: r2umin(*), r2atat(*)
      end function

      integer(c_int) function msg_create(params, ctx) bind(C, name='msg_create')
         import :: c_int, c_ptr, msg_params
         type(msg_params), intent(in) :: params
         type(c_ptr), intent(out) :: ctx
      end function

      subroutine msg_destroy(ctx) bind(C, name='msg_destroy')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine

      integer(c_int) function msg_set_table(ctx, ubuf, nbuf, iubuflow, nugrid, r2umin, r2atat) &
            bind(C, name='msg_set_table')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: ubuf(*)
         integer(c_int), value :: nbuf
         integer(c_int), intent(in) :: iubuflow(*), nugrid(*)
         real(c_double), intent(in) :: r2umin(*), r2atat(*)
      end function

      integer(c_int) function msg_set_particles(ctx, az, iptpn) bind(C, name='msg_set_particles')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: az(*)
         integer(c_int), intent(in) :: iptpn(*)
      end function

      integer(c_int) function msg_set_positions(ctx, ro) bind(C, name='msg_set_positions')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: ro(3,*)
      end function

      integer(c_int) function msg_ewald_setup(ctx) bind(C, name='msg_ewald_setup')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function msg_set_box(ctx, boxlen) bind(C, name='msg_set_box')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: boxlen(3)
      end function

      integer(c_int) function msg_nlist(ctx) bind(C, name='msg_nlist')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function msg_get_cell_list(ctx, nc, cell_start, cell_ip) bind(C, name='msg_get_cell_list')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: nc(3), cell_start(*), cell_ip(*)
      end function

      ! nmlIntList's inlist (nlist.F90:88-95,615-618): 0 skin test, > 0 fixed interval
      integer(c_int) function msg_nlist_interval(ctx, inlist) bind(C, name='msg_nlist_interval')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: inlist
      end function

      ! maxnneigh (nlist.F90:240-246); 0 sizes the list from the density
      integer(c_int) function msg_set_nlist_width(ctx, maxnneigh) bind(C, name='msg_set_nlist_width')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: maxnneigh
      end function

      ! NList(iBeforeMacrostep) (nlist.F90:593-595)
      integer(c_int) function msg_nlist_macrostep(ctx) bind(C, name='msg_nlist_macrostep')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_long_long) function msg_nlist_builds(ctx) bind(C, name='msg_nlist_builds')
         import :: c_long_long, c_ptr
         type(c_ptr), value :: ctx
      end function

      ! the Verlet list in MOLSIM's layout: kind 0 SetVListMD, 1 SetVListMC (nlist.F90:724-884)
      integer(c_int) function msg_get_verlet_list(ctx, kind, maxnneigh, nneighpn, jpnlist) &
            bind(C, name='msg_get_verlet_list')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: kind, maxnneigh
         integer(c_int), intent(out) :: nneighpn(*)
         integer(c_int), intent(out) :: jpnlist(maxnneigh,*)
      end function

      integer(c_int) function msg_utotal(ctx, u_twob, u_rec, u_tot, force, virial) &
            bind(C, name='msg_utotal')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: u_twob(0:*)
         real(c_double), intent(out) :: u_rec, u_tot
         real(c_double), intent(out) :: force(3,*)
         real(c_double), intent(out) :: virial
      end function

      integer(c_int) function msg_dutotal(ctx, nptm, ipnptm, rotm, lhsoverlap, du_twob, du_rec, du_tot) &
            bind(C, name='msg_dutotal')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nptm
         integer(c_int), intent(in) :: ipnptm(*)
         real(c_double), intent(in) :: rotm(3,*)
         integer(c_int), intent(out) :: lhsoverlap
         real(c_double), intent(out) :: du_twob(0:*)
         real(c_double), intent(out) :: du_rec, du_tot
      end function

      ! ntrial independent trials against the same state; outputs per trial
      integer(c_int) function msg_dutotal_batch(ctx, ntrial, ip, rotm, lhsoverlap, du_tot, du_rec) &
            bind(C, name='msg_dutotal_batch')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: ntrial
         integer(c_int), intent(in) :: ip(*)
         real(c_double), intent(in) :: rotm(3,*)
         integer(c_int), intent(out) :: lhsoverlap(*)
         real(c_double), intent(out) :: du_tot(*), du_rec(*)
      end function

      integer(c_int) function msg_commit(ctx, nptm, ipnptm, rotm) bind(C, name='msg_commit')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nptm
         integer(c_int), intent(in) :: ipnptm(*)
         real(c_double), intent(in) :: rotm(3,*)
      end function

      ! weak (titrating) charges, lweakcharge: laz(np) as 0/1, zat(npt) (energy.F90:7180-7276)
      integer(c_int) function msg_set_weak_charges(ctx, laz, zat) bind(C, name='msg_set_weak_charges')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(in) :: laz(*)
         real(c_double), intent(in) :: zat(*)
      end function

      ! ChargeChange trial (mc.F90:3683-3800): charge states of ipnptm(1:nptm) flip
      integer(c_int) function msg_dutotal_charge(ctx, nptm, ipnptm, lhsoverlap, du_twob, du_rec, du_tot) &
            bind(C, name='msg_dutotal_charge')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nptm
         integer(c_int), intent(in) :: ipnptm(*)
         integer(c_int), intent(out) :: lhsoverlap
         real(c_double), intent(out) :: du_twob(*), du_rec, du_tot
      end function

      ! MCUpdate of an accepted ChargeChange (mc.F90:3781-3785)
      integer(c_int) function msg_commit_charge(ctx, nptm, ipnptm) bind(C, name='msg_commit_charge')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nptm
         integer(c_int), intent(in) :: ipnptm(*)
      end function

      ! VelVer on the device (md.F90:252-410): masses per type, velocities rod(3,np)
      integer(c_int) function msg_md_init(ctx, mass, vel) bind(C, name='msg_md_init')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: mass(*)
         real(c_double), intent(in) :: vel(3,*)
      end function

      integer(c_int) function msg_md_run(ctx, nsteps, dt, log) bind(C, name='msg_md_run')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: nsteps
         real(c_double), value :: dt
         real(c_double), intent(out) :: log(2,*)
      end function

      integer(c_int) function msg_get_positions(ctx, ro) bind(C, name='msg_get_positions')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: ro(3,*)
      end function

      integer(c_int) function msg_get_velocities(ctx, vel) bind(C, name='msg_get_velocities')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: vel(3,*)
      end function

      ! MOLSIM's Random (mollib.F90:2689-2713) and MCPass of single-particle moves
      real(c_double) function msg_random(iseed, state, am) bind(C, name='msg_random')
         import :: c_int, c_double
         integer(c_int), intent(inout) :: iseed
         integer(c_int), intent(inout) :: state(2)
         real(c_double), intent(inout) :: am
      end function

      integer(c_int) function msg_mc_pass(ctx, ntrial, isamp, dtran, beta, iseed, state, am, &
                                          du_sum, naccept, events) bind(C, name='msg_mc_pass')
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: ntrial, isamp
         real(c_double), intent(in) :: dtran(*)
         real(c_double), value :: beta
         integer(c_int), intent(inout) :: iseed
         integer(c_int), intent(inout) :: state(2)
         real(c_double), intent(inout) :: am
         real(c_double), intent(out) :: du_sum
         integer(c_int), intent(out) :: naccept
         integer(c_int), intent(out) :: events(*)
      end function

   end interface

contains

   !> Stop with the library's message on a nonzero status (MOLSIM's Stop semantics).
   subroutine GpuCheck(rc, where)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: where
      character(len=256) :: text
      character(kind=c_char), pointer :: p(:)
      integer :: i
      if (rc == MSG_OK) return
      call c_f_pointer(msg_last_error(), p, [256])
      text = ' '
      do i = 1, 256
         if (p(i) == c_null_char) exit
         text(i:i) = p(i)
      end do
      write(*,'(a,a,a,i4,a,a)') 'molsim_gpu: ', where, ' failed with status ', rc, ': ', trim(text)
      stop 1
   end subroutine GpuCheck

end module MolsimGpu
