!> Fortran caller of the C-ABI through the MolsimGpu module, in the way a MOLSIM build calls
!! it: reads a system (positions, charges, types, box, Ewald parameters) from a text file,
!! builds the table (PotTwoBodyTab semantics), runs UTotal, a few DUTotal trials (one at a time
!! and batched), reads the cell list, runs one MCUpdate and a short MCPass, and prints the
!! results for the caller to compare.
!!
!! Input file (free format): np npt / boxlen(3) / rcut alpha epsi4pi / nmesh order lsurf /
!! form lewald / z(npt) / sig(npt) / eps(npt) / rad(npt) / (ro(1:3,ip), az(ip), iptpn(ip), ip=1,np)
!! / ntrial / (ip, rotm(1:3), itrial=1,ntrial)
program drop_in_driver

   use MolsimGpu
   implicit none

   type(msg_params) :: prm
   type(msg_pot) :: pot
   type(c_ptr) :: ctx
   character(len=512) :: fin
   integer :: ip, npt, i, ntrial, it
   integer(c_int) :: rc, nbuf, lsurf, form, lewald, ovl, nmesh, order
   integer(c_int), allocatable :: iptpn(:), iubuflow(:), nugrid(:), iptrial(:)
   real(c_double), allocatable :: ro(:,:), az(:), force(:,:), ubuf(:), r2umin(:), r2atat(:), rtrial(:,:)
   real(c_double), allocatable :: u_twob(:), du_twob(:)
   real(c_double) :: boxlen(3), rcut, alpha, epsi4pi, u_rec, u_tot, virial, du_rec, du_tot
   integer(c_int) :: iseed, rngstate(2), naccept
   integer(c_int), allocatable :: events(:), ovlb(:), cell_start(:), cell_ip(:)
   integer(c_int) :: ncl(3)
   real(c_double), allocatable :: dutb(:), durb(:)
   real(c_double) :: am, dusum, dtran(8)

   call get_command_argument(1, fin)
   open(10, file=trim(fin), status='old')
   read(10,*) prm%np, npt
   read(10,*) boxlen
   read(10,*) rcut, alpha, epsi4pi
   read(10,*) nmesh, order, lsurf
   read(10,*) form, lewald
   pot%npt = npt
   pot%form = form
   pot%lewald = lewald
   pot%pad = 0
   pot%alpha = alpha
   pot%rcut = rcut
   pot%epsi4pi = epsi4pi
   pot%utoltab = 1d-5
   pot%ftoltab = 1d-5
   pot%umaxtab = 2d4
   pot%fmaxtab = 2d4
   pot%r2uminin = 0.1d0
   pot%z = 0; pot%sig = 0; pot%eps = 0; pot%rad = 0
   read(10,*) pot%z(1:npt)
   read(10,*) pot%sig(1:npt)
   read(10,*) pot%eps(1:npt)
   read(10,*) pot%rad(1:npt)
   allocate(ro(3,prm%np), az(prm%np), iptpn(prm%np), force(3,prm%np))
   do ip = 1, prm%np
      read(10,*) ro(1:3,ip), az(ip), iptpn(ip)
   end do
   read(10,*) ntrial
   allocate(iptrial(ntrial), rtrial(3,ntrial))
   do it = 1, ntrial
      read(10,*) iptrial(it), rtrial(1:3,it)
   end do
   close(10)

   ! PotTwoBodyTab: the table MOLSIM would hold
   allocate(ubuf(200000), iubuflow(npt*(npt+1)/2), nugrid(npt*(npt+1)/2))
   allocate(r2umin(npt*(npt+1)/2), r2atat(npt*(npt+1)/2))
   nbuf = msg_build_table(pot, ubuf, 200000_c_int, iubuflow, nugrid, r2umin, r2atat)
   if (nbuf < 0) call GpuCheck(nbuf, 'msg_build_table')

   prm%npt = npt
   prm%boxlen = boxlen
   prm%rcut = rcut
   prm%drnlist = 0d0
   prm%epsi4pi = epsi4pi
   prm%list_mode = MSG_LIST_CELL
   prm%ewald_mode = MSG_EWALD_SPM
   prm%lsurf = lsurf
   prm%alpha = alpha
   prm%ncut = 0
   prm%ncutregion = 0
   prm%nmesh = nmesh
   prm%order = order
   prm%precision = MSG_PREC_MIXED
   prm%force_scale = 1d0

   call GpuCheck(msg_create(prm, ctx), 'msg_create')
   call GpuCheck(msg_set_table(ctx, ubuf, nbuf, iubuflow, nugrid, r2umin, r2atat), 'msg_set_table')
   call GpuCheck(msg_set_particles(ctx, az, iptpn), 'msg_set_particles')
   call GpuCheck(msg_set_positions(ctx, ro), 'msg_set_positions')
   call GpuCheck(msg_ewald_setup(ctx), 'msg_ewald_setup')

   allocate(u_twob(0:npt*(npt+1)/2), du_twob(0:npt*(npt+1)/2))
   call GpuCheck(msg_utotal(ctx, u_twob, u_rec, u_tot, force, virial), 'msg_utotal')
   write(*,'(a,es24.16)') 'u_tot ', u_tot
   write(*,'(a,es24.16)') 'u_rec ', u_rec
   write(*,'(a,es24.16)') 'u_twob0 ', u_twob(0)
   write(*,'(a,es24.16)') 'virial ', virial
   do ip = 1, prm%np
      write(*,'(a,i8,3es24.16)') 'force ', ip, force(1:3,ip)
   end do

   do it = 1, ntrial
      call GpuCheck(msg_dutotal(ctx, 1_c_int, iptrial(it:it), rtrial(:,it:it), ovl, du_twob, du_rec, du_tot), &
                    'msg_dutotal')
      write(*,'(a,i6,i3,es24.16)') 'dutotal ', it, ovl, du_tot
   end do

   ! the same trials as one batched call
   allocate(ovlb(ntrial), dutb(ntrial), durb(ntrial))
   call GpuCheck(msg_dutotal_batch(ctx, int(ntrial, c_int), iptrial, rtrial, ovlb, dutb, durb), 'msg_dutotal_batch')
   do it = 1, ntrial
      write(*,'(a,i6,i3,es24.16)') 'dutotal_batch ', it, ovlb(it), dutb(it)
   end do

   ! the cell list MOLSIM's linked lists would hold
   allocate(cell_start(prm%np + 100000), cell_ip(prm%np))
   call GpuCheck(msg_get_cell_list(ctx, ncl, cell_start, cell_ip), 'msg_get_cell_list')
   write(*,'(a,3i6,i10)') 'cells ', ncl, sum(int(cell_ip, 8) * [(int(i, 8), i = 1, prm%np)])

   ! MCUpdate of the first trial, then UTotal again
   call GpuCheck(msg_commit(ctx, 1_c_int, iptrial(1:1), rtrial(:,1:1)), 'msg_commit')
   call GpuCheck(msg_utotal(ctx, u_twob, u_rec, u_tot, force, virial), 'msg_utotal')
   write(*,'(a,es24.16)') 'u_tot_after ', u_tot

   ! MCPass of 50 sequential single-particle moves with MOLSIM's generator
   iseed = -4711
   rngstate = -1
   am = 0d0
   dtran = 2d0
   allocate(events(50))
   call GpuCheck(msg_mc_pass(ctx, 50_c_int, 0_c_int, dtran, 1d3 / (8.314472d0 * 298d0), iseed, &
                             rngstate, am, dusum, naccept, events), 'msg_mc_pass')
   write(*,'(a,i6,es24.16)') 'mc_pass ', naccept, dusum
   write(*,'(a,i12,2i12)') 'rng_state ', iseed, rngstate

   call msg_destroy(ctx)

end program drop_in_driver
