! Writes the frames of the reference's write_model (src/read_input.f90:403-421) with its
! own edit descriptors, for values read from a stream file: the byte-exact target of the
! library's frame formatter.
program frames_ref
  implicit none
  integer :: np, ntypes, step, i, u
  character(len=8), allocatable :: names(:)
  integer, allocatable :: typ(:)
  real(8), allocatable :: pos(:,:), vel(:,:), pene(:)
  real(8) :: box(3)
  open(newunit=u, file='frame_in.bin', access='stream', form='unformatted', status='old')
  read(u) np, ntypes, step
  allocate(names(ntypes), typ(np), pos(3,np), vel(3,np), pene(np))
  read(u) names, typ, pos, vel, pene, box
  close(u)
  open(newunit=u, file='frame_ref.xyz', status='replace')
  write(u,'(i8)') np
  write(u,'(A)')
  do i = 1, np
    write(u,'(A4,x,7EN14.4)') names(typ(i)), pos(1:3,i), vel(1:3,i), pene(i)
  enddo
  close(u)
  open(newunit=u, file='frame_ref.dump', status='replace')
  write(u,'(A)') 'ITEM: TIMESTEP'
  write(u,'(i7)') step
  write(u,'(A)') 'ITEM: NUMBER OF ATOMS'
  write(u,'(i7)') np
  write(u,'(A)') 'ITEM: BOX BOUNDS pp pp pp'
  write(u,'(2f12.4)') 0.0, box(1)
  write(u,'(2f12.4)') 0.0, box(2)
  write(u,'(2f12.4)') 0.0, box(3)
  write(u,'(A)') 'ITEM: ATOMS id type xs ys zs'
  do i = 1, np
    write(u,'(i7,x,i2,3f9.5)') i, typ(i), real(pos(1:3,i))+0.0
  enddo
  close(u)
end program
