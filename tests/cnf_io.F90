! Writes and reads MOLSIM .cnf files with the I/O lists of IOCnf (Src/molsim.F90:876-884,
! 895-903, 936-941): the Fortran side of the tests of molsim_amd.cnf, compiled by the test.
!   cnf_io write <path>   writes a 5-particle file with weak-charge and velocity records
!   cnf_io read <path>    reads such a file and prints every value
program cnf_io
   implicit none
   character(len=512) :: mode, path
   integer(4) :: istep1, iseed, ix, iy, np, iaux, ip, i
   real(8) :: am, boxlen(3), raux
   logical :: laux
   real(8), allocatable :: ro(:,:), qua(:,:), rod(:,:), quad(:,:)
   logical, allocatable :: laz(:)

   call get_command_argument(1, mode)
   call get_command_argument(2, path)
   if (trim(mode) == 'write') then
      np = 5
      allocate(ro(3,np), qua(0:3,np), rod(3,np), quad(0:3,np), laz(np))
      istep1 = 42
      iseed = 4712
      am = 4.6566125955216364d-10
      ix = -993160733
      iy = 1260042462
      boxlen = [10.5d0, 11.25d0, 12.125d0]
      do ip = 1, np
         do i = 1, 3
            ro(i,ip) = ip + 0.125d0*i - 3.0d0
            rod(i,ip) = -ip - 0.5d0*i
         end do
         qua(0:3,ip) = [0.5d0, -0.5d0, 0.5d0, -0.5d0] * ip / ip
         quad(0:3,ip) = [0.0d0, 1.0d0*ip, -2.0d0, 0.25d0]
         laz(ip) = mod(ip, 2) == 0
      end do
      open(10, file=trim(path), form='unformatted', access='sequential', status='replace')
      write(10) istep1, iseed, am, ix, iy, boxlen
      write(10) np, 0, 0.0d0, 0.0d0, .false., .false.
      write(10) (ro(1:3,ip), qua(0:3,ip), ip = 1, np)
      write(10) laz(1:np)
      write(10) (rod(1:3,ip), quad(0:3,ip), ip = 1, np)
      close(10)
   else
      open(10, file=trim(path), form='unformatted', access='sequential', status='old')
      read(10) istep1, iseed, am, ix, iy, boxlen
      read(10) np, iaux, raux, raux, laux, laux
      allocate(ro(3,np), qua(0:3,np), rod(3,np), quad(0:3,np), laz(np))
      read(10) (ro(1:3,ip), qua(0:3,ip), ip = 1, np)
      read(10) laz(1:np)
      read(10) (rod(1:3,ip), quad(0:3,ip), ip = 1, np)
      close(10)
      print '(i0,1x,i0,1x,es26.17e3,1x,i0,1x,i0)', istep1, iseed, am, ix, iy
      print '(3es26.17e3)', boxlen
      print '(i0)', np
      do ip = 1, np
         print '(14es26.17e3,1x,l1)', ro(1:3,ip), qua(0:3,ip), rod(1:3,ip), quad(0:3,ip), laz(ip)
      end do
   end if
end program cnf_io
