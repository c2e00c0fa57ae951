! ref_harness.f90 -- drives the compiled Ndyn reference to write golden vectors.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
! src/{mods,domtype,potentials,read_input,nd}.f90 where they lie (never copied),
! renames the reference's main program symbol in its nd.o, and links this harness
! in its place.  The harness sets the reference's module state directly (so the
! thread-dependent Velocity RNG, src/nd.f90:1500-1624, is bypassed), calls the
! reference routines, and dumps their results at full precision.
!
! Input (stream binary, little endian):
!   int32  mode, np, ntypes, npairs, ptsize, neimax, nsteps, outsteps, step0,
!          renei, renei_freq, isokinetic, ke_freq, isobaric, p_freq,
!          isoenergetic, e_freq, nthreads
!   real64 box(3), dt, skin, ptrmin, ketarget, keparm, ptarget(3), pparm, eparm
!   real64 mass(ntypes); int32 pairmat(ntypes,ntypes); int32 kind(npairs)
!   real64 pairparams(4,npairs); int32 typ(np); real64 pos(3,np), vel(3,np), acc(3,np)
!   int32  boxload, have_ipos; real64 boxvel(3)
!   (have_ipos /= 0) real64 ipos(3,np), initbox(3)
!   int32 nhold; (nhold > 0) int32 hold(np), bcdisp(nhold), bcdir(3,nhold);
!   real64 bcparm(3,nhold); real64 pairchargeprod(npairs)
! Pair kinds: 1 morse, 2 buck, 3 LJ, 5 spring, 6 sine, 7 cosine, 8 soft, 9 John,
! 10 MorseC, 11 BuckC, 12 LJC.
! Modes: 1 tables (buildPairTables), 2 half list (buildCellList),
!        3 forces (buildCellList + the force routine dynamics would call: compute0,
!        compute1 for an asymmetric PairMat, compute2/compute3 with spring pairs),
!        4 'dynamics' run.
program ref_harness
  use omp_lib
  use mostCommon
  use interpot
  use control
  use listData
  use groupStuff
  implicit none
  character(len=1024) :: fin, fout
  integer :: mode, nthr, i
  integer :: iv(18), iv2(2), nh
  integer, allocatable :: ibc(:), ibcdir(:,:)
  real(8) :: dv(13)
  integer, allocatable :: kinds(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(20, file=trim(fin), access='stream', form='unformatted', status='old')
  read(20) iv
  mode = iv(1); np = iv(2); Ntypes = iv(3); NPairs = iv(4); PTsize = iv(5)
  NeiMax = iv(6); step_num = iv(7); OutSteps = iv(8); step = iv(9)
  ReNei = iv(10) /= 0; ReNeiFreq = iv(11)
  isokinetic = iv(12) /= 0; KEfreq = iv(13)
  isobaric = iv(14) /= 0; Pfreq = iv(15)
  isoenergetic = iv(16) /= 0; Efreq = iv(17)
  nthr = iv(18)
  read(20) dv
  box = dv(1:3); dt = dv(4); skin = dv(5); PTrMin = dv(6)
  KEtarget = dv(7); KEparm = dv(8); Ptarget = dv(9:11); Pparm = dv(12); Eparm = dv(13)

  call omp_set_num_threads(nthr)
  thread_num = nthr
  proc_num = omp_get_num_procs()

  allocate(mass(Ntypes), typeName(Ntypes), charge(Ntypes), NatomsPerType(0:Ntypes))
  read(20) mass
  charge = 0.d0
  do i = 1, Ntypes
    write(typeName(i), '(A,i0)') 'T', i
  enddo
  allocate(PairMat(Ntypes,Ntypes), PairName(0:NPairs), PairParams(4,NPairs), &
           PairChargeProd(NPairs), PTdr2(NPairs), PTrc2(NPairs), kinds(NPairs))
  read(20) PairMat
  read(20) kinds
  read(20) PairParams
  PairName = 'N/A'; PairChargeProd = 0.d0; PTdr2 = 0.d0; PTrc2 = 0.d0
  do i = 1, NPairs
    select case (kinds(i))
      case (1); PairName(i) = 'morse'
      case (2); PairName(i) = 'buck'
      case (3); PairName(i) = 'LJ'
      case (5); PairName(i) = 'spring'
      case (6); PairName(i) = 'sine'
      case (7); PairName(i) = 'cosine'
      case (8); PairName(i) = 'soft'
      case (9); PairName(i) = 'John'
      case (10); PairName(i) = 'MorseC'
      case (11); PairName(i) = 'BuckC'
      case (12); PairName(i) = 'LJC'
      case default; stop 'harness: unknown pair kind'
    end select
  enddo
  ! the deck's 'asym' flag (src/read_input.f90:130-133): some PairMat(a,b) /= PairMat(b,a)
  asympots = any(PairMat /= transpose(PairMat)); springpots = any(kinds == 5)
  potOuts = .false.

  allocate(typ(np), pos(3,np), vel(3,np), acc(3,np), force(3,np), pene(np))
  read(20) typ
  read(20) pos
  read(20) vel
  read(20) acc
  read(20) iv2
  read(20) boxVel
  if (iv2(2) /= 0) then
    allocate(ipos(3,np))
    read(20) ipos
    read(20) initBox
  endif
  read(20) nh
  close(20)
  call read_charge_products()
  force = 0.d0; pene = 0.d0
  grpBuff = 100; holdBuff = 100; Ngrp = 0; Nhold = 0
  allocate(grp(np), groupName(grpBuff), hold(np), holdName(holdBuff), &
           bcTyp(holdBuff), bcParm(3,holdBuff), bcDir(3,holdBuff))
  grp = 0; hold = 0; bcTyp = 'X'; bcParm = 0.d0; bcDir = .false.
  if (nh > 0) then
    ! Hold groups and their Disp conditions (src/read_input.f90:177-190)
    allocate(ibc(nh), ibcdir(3,nh))
    open(20, file=trim(fin), access='stream', form='unformatted', status='old')
    call skip_to_holds()
    read(20) hold
    read(20) ibc
    read(20) ibcdir
    read(20) bcParm(1:3,1:nh)
    close(20)
    Nhold = nh
    do i = 1, nh
      if (ibc(i) /= 0) bcTyp(i) = 'd'
      bcDir(1:3,i) = ibcdir(1:3,i) /= 0
    enddo
  endif
  boxLoad = iv2(1) /= 0
  Model = 'NA'; ModelInType = 'NA'; ModelOutType = 'dump'; ConfOuts = 0
  potential = 0.d0; kinetic = 0.d0; virial = 0.d0; e0 = 0.d0

  call buildPairTables

  open(21, file=trim(fout), access='stream', form='unformatted', status='replace')
  select case (mode)
  case (1)
    write(21) PTrc2, PTdr2, PPairTable, FPairTable
  case (2, 3)
    ! the prologue of 'dynamics' (src/nd.f90:77-107)
    if (skin < 0.d0) skin = rcut*0.2d0
    MCELL(1) = int(box(1)/(rcut+skin))
    MCELL(2) = int(box(2)/(rcut+skin))
    MCELL(3) = int(box(3)/(rcut+skin))
    NCELL = MCELL(1)*MCELL(2)*MCELL(3)
    if (NeiMax == 0) then
      NeiMax = int(4./3.*PI*((rcut+skin)*1.1)**3*np/(box(1)*box(2)*box(3)))
    endif
    allocate(HEAD(NCELL), VNL(1:NeiMax+1,1:np), List(np), dispList(3,np), &
             ftmp(1:3,thread_num*np))
    skinsq = skin*skin
    call buildCellList(.false.)
    if (mode == 2) then
      write(21) MCELL, NeiMax
      write(21) VNL
    else
      if (asympots .and. springpots) then
        call compute3(.true., .true.)
      else if (springpots) then
        call compute2(.true., .true.)
      else if (asympots) then
        call compute1(.true., .true.)
      else
        call compute0(.true., .true.)
      endif
      write(21) force, potential, virial
    endif
  case (4)
    call dynamics
    write(21) pos, vel, acc, force, potential, kinetic, virial, box, e0, &
              dble(NlistUpdates), dble(step)
  case default
    stop 'harness: unknown mode'
  end select
  close(21)
contains
  ! PairChargeProd, the input's last block
  subroutine read_charge_products()
    integer(8) :: off
    off = 4*18 + 8*13 + 8*Ntypes + 4*Ntypes*Ntypes + 4*NPairs + 8*4*NPairs + 4*np + 3*8*3*np &
          + 4*2 + 8*3 + 4
    if (iv2(2) /= 0) off = off + 8*3*np + 8*3
    if (nh > 0) off = off + 4*np + 4*nh + 4*3*nh + 8*3*nh
    open(20, file=trim(fin), access='stream', form='unformatted', status='old')
    read(20, pos=off+1) PairChargeProd
    close(20)
  end subroutine read_charge_products

  ! position the input at the hold block: after the fixed header, the per-type and per-pair
  ! arrays, the state, the BoxVel block and the optional initial positions
  subroutine skip_to_holds()
    integer(8) :: off
    off = 4*18 + 8*13 + 8*Ntypes + 4*Ntypes*Ntypes + 4*NPairs + 8*4*NPairs + 4*np + 3*8*3*np &
          + 4*2 + 8*3 + 4
    if (iv2(2) /= 0) off = off + 8*3*np + 8*3
    read(20, pos=off+1)
  end subroutine skip_to_holds
end program ref_harness
