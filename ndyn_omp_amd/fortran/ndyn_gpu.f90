! ndyn_gpu.f90 -- Ndyn deck driver over the MI355X engine (C ABI via iso_c_binding).
!
! Reads the same input deck on stdin as the reference (src/read_input.f90:60-206: START,
! Types, Pairs, Model, BoxSize, NTotal, Seed, Velocity, Skin, NeiMax, ReNeighbor,
! Isokinetic/Isobaric/Isoenergetic and their Aniso* forms, BoxVel/StopBoxVel, OutFreq,
! ConfOutFreq, StepNumber, WritePotentialTable, Exec, Print, Run, END), builds the same
! potential tables (src/potentials.f90:2-56, through ndg_pair_table), and writes the same
! thermo lines (src/read_input.f90:433-446) and trajectory files out.xyz / out.dump
! (src/read_input.f90:394-431).  The MD step runs on the GPU ('dynamics',
! src/nd.f90:60-199, through ndg_run_begin / ndg_run_steps / ndg_run_end).
! Three directives are added: 'Precision mixed|fp64' selects the engine path,
! 'ListTrigger compat|skin' the neighbour-list rebuild criterion, and
! 'FrameWriter engine|fortran' who formats the trajectory frames: the library's writer
! thread (default; same bytes, formatted while the device keeps stepping) or write_model.
! Asymmetric pairs ('asym', the reference's compute1), spring pairs (compute2/3), InitPos
! and the Hold/Disp boundary conditions run on the GPU as well; Group/Hold select atoms by
! type or geometric domain (src/nd.f90:1625-1730, src/domtype.f90) and ChangeType and
! Velocity act on groups, on the host between Runs.
program ndyn_gpu
  use iso_c_binding
  use ndg_iface
  implicit none

  ! system
  integer :: np = 1000, ntypes = 0, npairs = 0, ptsize = 1000
  real(8) :: ptrmin = 0.25d0, box(3) = 10.d0
  integer, allocatable :: typ(:), pairmat(:,:), pairkind(:)
  real(8), allocatable :: pos(:,:), vel(:,:), acc(:,:), pene(:), frc(:,:)
  real(8), allocatable :: mass(:), pairparams(:,:), ptab(:,:), ftab(:,:), rc2(:), dr2(:)
  real(8), allocatable :: charge(:), pairq(:)
  character(len=8), allocatable :: typename(:), pairname(:)
  ! control (defaults of src/read_input.f90:18-41)
  real(8) :: dt = 0.0001d0, skin = -1.d0
  integer :: step = 0, step_num = 400, outsteps = 40, confouts = 0, neimax = 0
  logical :: renei = .true., isokin = .false., isobar = .false., isoen = .false., boxload = .false.
  integer :: renei_freq = 1, kefreq = 1, pfreq = 1, efreq = 1
  real(8) :: ketarget = 0.d0, keparm = 0.d0, ptarget(3) = 0.d0, pparm = 0.d0, eparm = 0.d0
  real(8) :: boxvel(3) = 0.d0
  character(len=10) :: outtype = 'NA'
  logical :: potouts = .false.
  integer :: path = NDG_PATH_MIXED, trigger = NDG_TRIGGER_COMPAT
  ! frames formatted by the library's writer thread (engine) or by write_model (fortran)
  logical :: engine_frames = .true.
  logical :: initpos_pending = .false.
  ! groups and hold groups (module groupStuff, src/mods.f90:11-17)
  integer :: ngrp = 0, nhold = 0, grpbuff = 100, holdbuff = 100
  integer, allocatable :: grp(:), hold(:), bcdisp(:), bcdir(:,:)
  character(len=8), allocatable :: groupname(:), holdname(:)
  real(8), allocatable :: bcparm(:,:)
  logical :: holds_dirty = .false.
  ! a geometric domain (type domain, src/domtype.f90:10-16)
  type :: domain_t
    character(len=3) :: ct = '   '
    logical :: inside = .true.
    integer :: axis = 0
    real(4) :: parm(8) = 0.0
  end type domain_t
  ! engine handle
  type(c_ptr) :: h = c_null_ptr
  logical :: tables_dirty = .true., state_dirty = .true., have_model = .false.
  ! parsing
  character(len=256) :: line, w(16)
  integer :: nw, ios, i, j, ln
  real(8) :: rtmp

  write(6,'(a)') ''
  write(6,'(a)') '#  N D Y N  (MI355X engine)'
  write(6,'(a)') ''
  write(6,'(A)') "#  type 'START' to begin specifying directives"
  call flush(6)

  do   ! wait for START (src/read_input.f90:48-57)
    read(5,'(A)',iostat=ios) line
    if (ios /= 0) call die('#  read_input: ERROR: premature end of input file. No END directive found')
    call split(line, w, nw)
    if (nw < 1) cycle
    if (trim(w(1)) == 'START') exit
    if (any(trim(w(1)) == [character(len=5) :: 'END', 'DONE', 'EXIT', 'quit', 'abort'])) stop
  enddo

  ln = 0
  do
    read(5,'(A)',iostat=ios) line
    if (ios /= 0) call die('#  read_input: ERROR: premature end of input file. No END directive found')
    ln = ln + 1
    call strip_comment(line)
    call split(line, w, nw)
    if (nw < 1) cycle
    select case (trim(w(1)))
    case ('help')
      call help
    case ('BoxSize')
      do i = 1, 3
        read(w(1+i),*) box(i)
      enddo
      state_dirty = .true.
    case ('Model')
      call read_model(trim(w(2)), merge(trim(w(3)), 'xyz ', nw >= 3))
    case ('Seed')
      call set_seed(w(2))
    case ('NTotal')
      read(w(2),*) np
      call random_model
    case ('Velocity')
      read(w(2),*) rtmp
      if (nw >= 3) then
        call velocities(rtmp, trim(w(3)))
      else
        call velocities(rtmp, 'N/A')
      endif
    case ('ReNeibour', 'ReNeighbor')
      read(w(2),*) renei
      if (nw >= 3) read(w(3),*) renei_freq
    case ('Skin')
      read(w(2),*) skin
    case ('NeiMax')
      read(w(2),*) neimax
    case ('Types')
      read(w(2),*) ntypes
      if (allocated(mass)) deallocate(mass, typename, charge)
      allocate(mass(ntypes), typename(ntypes), charge(ntypes))
      charge = 0.d0
      do i = 1, ntypes
        read(5,'(A)',iostat=ios) line
        if (ios /= 0) call die('#  read_input: ERROR: premature end of input file. No END directive found')
        call strip_comment(line)
        call split(line, w, nw)
        typename(i) = trim(w(1))
        read(w(2),*) mass(i)
        ! name mass [charge] (src/read_input.f90:100-108)
        if (nw >= 3) read(w(3),*) charge(i)
      enddo
      tables_dirty = .true.
    case ('WritePotentialTable')
      potouts = .true.
    case ('Pairs')
      read(w(2),*) npairs
      if (nw >= 3) read(w(3),*) ptrmin
      if (nw >= 4) read(w(4),*) ptsize
      call read_pairs
    case ('Isokinetic')
      isokin = .true.
      read(w(2),*) ketarget
      read(w(3),*) keparm
      if (nw >= 4) read(w(4),*) kefreq
    case ('Anisokinetic')
      isokin = .false.
    case ('Isoenergetic')
      isoen = .true.
      read(w(2),*) eparm
      if (nw >= 3) read(w(3),*) efreq
    case ('Anisoenergetic')
      isoen = .false.
    case ('Isobaric')
      isobar = .true.
      do i = 1, 3
        read(w(1+i),*) ptarget(i)
      enddo
      read(w(5),*) pparm
      if (nw >= 6) read(w(6),*) pfreq
    case ('Anisobaric')
      isobar = .false.
    case ('BoxVel')
      boxload = .true.
      do i = 1, 3
        read(w(1+i),*) boxvel(i)
      enddo
    case ('StopBoxVel')
      boxload = .false.
    case ('Exec')
      call execute_command_line(trim(adjustl(line(5:))))
    case ('Print')
      write(6,'(A)') '# '//trim(adjustl(line(6:)))
    case ('OutFreq')
      read(w(2),*) outsteps
    case ('StepNumber')
      read(w(2),*) step
    case ('ConfOutFreq')
      read(w(2),*) confouts
      if (nw >= 3) outtype = trim(w(3))
    case ('Precision')
      if (trim(w(2)) == 'fp64') then
        path = NDG_PATH_FP64
      else
        path = NDG_PATH_MIXED
      endif
    case ('FrameWriter')
      engine_frames = trim(w(2)) /= 'fortran'
    case ('ListTrigger')
      if (trim(w(2)) == 'skin') then
        trigger = NDG_TRIGGER_SKIN
      else
        trigger = NDG_TRIGGER_COMPAT
      endif
    case ('Run')
      read(w(2),*) step_num
      if (nw >= 3) read(w(3),*) dt
      call dynamics
    case ('InitPos')
      ! initPositionSet (src/read_input.f90:165, src/nd.f90:1731-1763), applied to the
      ! state the next Run starts from
      if (allocated(pairkind)) then
        if (any(iand(pairkind, not(NDG_ESHIFT)) == NDG_SPRING)) initpos_pending = .true.
      endif
    case ('GroupBufferSize')
      read(w(2),*) grpbuff
      call alloc_groups(.true.)
    case ('HoldBufferSize')
      read(w(2),*) holdbuff
      call alloc_groups(.true.)
    case ('Group')   ! src/read_input.f90:171-174
      call alloc_groups(.false.)
      if (ngrp >= grpbuff) call die('ERROR: more groups than GroupBufferSize')
      ngrp = ngrp + 1
      groupname(ngrp) = trim(w(2))
      call set_group('g', ngrp)
    case ('ClearGroups')
      ngrp = 0
      if (allocated(grp)) grp = 0
    case ('Hold')     ! src/read_input.f90:177-180
      call alloc_groups(.false.)
      if (nhold >= holdbuff) call die('ERROR: more hold groups than HoldBufferSize')
      nhold = nhold + 1
      holdname(nhold) = trim(w(2))
      call set_group('h', nhold)
      holds_dirty = .true.
    case ('ClearHolds')
      ! the Disp settings of the cleared groups stay with their indices, as in the reference
      nhold = 0
      if (allocated(hold)) hold = 0
      holds_dirty = .true.
    case ('Disp')     ! src/read_input.f90:184-190
      call alloc_groups(.false.)
      i = hold_id(w(2))
      if (i == 0) call die('ERROR: Disp: unknown hold group '//trim(w(2)))
      bcdisp(i) = 1
      do j = 1, 3
        if (trim(w(2+j)) /= 'F') then
          bcdir(j,i) = 1
          read(w(2+j),*) bcparm(j,i)
        endif
      enddo
      holds_dirty = .true.
    case ('ChangeType')   ! setType, src/nd.f90:1700-1730
      call change_type(w(2), w(3))
    case ('END', 'DONE', 'end', 'done')
      exit
    case ('EXIT', 'exit', 'quit', 'abort')
      stop
    case default
      write(0,'(A,i4)') '#  read_input: ERROR: Unknown directive: '//trim(w(1))//' line:', ln
      call flush(0)
    end select
  enddo
  if (c_associated(h)) i = ndg_destroy(h)
  write(6,'(a)') '#  Ndyn:  Normal end of execution.'

contains

  subroutine die(msg)
    character(*), intent(in) :: msg
    write(0,'(A)') msg
    call flush(0)
    stop 1
  end subroutine die

  subroutine check(ierr)
    integer(c_int), intent(in) :: ierr
    select case (ierr)
    case (NDG_OK)
      return
    case (NDG_ERR_NEIMAX)
      call die('too many neighbours in list')
    case (NDG_ERR_OUTSIDE)
      call die('Particle outside of box')
    case default
      call die('engine error: '//trim(ndg_error_text()))
    end select
  end subroutine check

  subroutine strip_comment(s)
    character(*), intent(inout) :: s
    integer :: k
    k = index(s, '#')
    if (k > 0) s(k:) = ' '
  end subroutine strip_comment

  subroutine split(s, words, n)
    character(*), intent(in) :: s
    character(len=256), intent(out) :: words(:)
    integer, intent(out) :: n
    integer :: k, b
    logical :: inw
    n = 0
    inw = .false.
    words = ' '
    b = 1
    do k = 1, len_trim(s) + 1
      if (k <= len_trim(s)) then
        if (s(k:k) /= ' ' .and. s(k:k) /= char(9)) then
          if (.not. inw) then
            inw = .true.
            b = k
          endif
          cycle
        endif
      endif
      if (inw) then
        inw = .false.
        if (n < size(words)) then
          n = n + 1
          words(n) = s(b:k-1)
        endif
      endif
    enddo
  end subroutine split

  integer function type_id(name)
    character(*), intent(in) :: name
    integer :: k
    type_id = 0
    do k = 1, ntypes
      if (trim(name) == trim(typename(k))) type_id = k
    enddo
  end function type_id

  subroutine alloc_atoms
    if (allocated(pos)) deallocate(pos, vel, acc, typ, pene, frc)
    allocate(pos(3,np), vel(3,np), acc(3,np), typ(np), pene(np), frc(3,np))
    pos = 0.d0; vel = 0.d0; acc = 0.d0; typ = 1; pene = 0.d0; frc = 0.d0
    have_model = .true.
    state_dirty = .true.
    if (allocated(grp)) deallocate(grp, hold)
    allocate(grp(np), hold(np))
    grp = 0; hold = 0
    ngrp = 0; nhold = 0
    holds_dirty = .true.
  end subroutine alloc_atoms

  ! group tables (src/read_input.f90:292-296); fresh = .true. reallocates at a new size
  subroutine alloc_groups(fresh)
    logical, intent(in) :: fresh
    if (allocated(groupname) .and. .not. fresh) return
    if (allocated(groupname)) deallocate(groupname, holdname, bcdisp, bcdir, bcparm)
    allocate(groupname(grpbuff), holdname(holdbuff), bcdisp(holdbuff), bcdir(3,holdbuff), &
             bcparm(3,holdbuff))
    bcdisp = 0; bcdir = 0; bcparm = 0.d0
    ngrp = min(ngrp, grpbuff); nhold = min(nhold, holdbuff)
  end subroutine alloc_groups

  integer function group_id(name)
    character(*), intent(in) :: name
    integer :: k
    group_id = 0
    do k = 1, ngrp
      if (trim(name) == trim(groupname(k))) then
        group_id = k
        return
      endif
    enddo
  end function group_id

  integer function hold_id(name)
    character(*), intent(in) :: name
    integer :: k
    hold_id = 0
    do k = 1, nhold
      if (trim(name) == trim(holdname(k))) then
        hold_id = k
        return
      endif
    enddo
  end function hold_id

  ! setGroup (src/nd.f90:1625-1699): words 3.. of the directive are 'typ N', 'dom ...' or
  ! 'all'; g = 'g' sets grp, 'h' sets hold, to id
  subroutine set_group(g, id)
    character, intent(in) :: g
    integer, intent(in) :: id
    integer :: k, cht
    type(domain_t) :: dom
    if (.not. have_model) call die('ERROR: Group/Hold needs a model')
    select case (w(3)(1:3))
    case ('typ')
      read(w(4),*) cht
      do k = 1, np
        if (typ(k) == cht) call put(k)
      enddo
    case ('dom')
      call read_domain(dom)
      do k = 1, np
        if (in_domain(real(pos(1:3,k)), dom)) call put(k)
      enddo
    case ('all')
      if (g == 'g') grp = 0
      if (g == 'h') hold = 0
    case default
      call die('ERROR: Group/Hold: expected typ, dom or all')
    end select
  contains
    subroutine put(k)
      integer, intent(in) :: k
      if (g == 'g') grp(k) = id
      if (g == 'h') hold(k) = id
    end subroutine put
  end subroutine set_group

  ! readDomain (src/domtype.f90:24-86): words 4.. = type, inside flag, [axis,] parameters
  subroutine read_domain(dom)
    type(domain_t), intent(out) :: dom
    integer :: k, np_, first
    dom%ct = trim(w(4))
    read(w(5),*) dom%inside
    select case (dom%ct)
    case ('rec'); np_ = 6; first = 6
    case ('per'); np_ = 7; first = 6
    case ('sph'); np_ = 4; first = 6
    case ('ell'); np_ = 4; first = 7
    case ('tub'); np_ = 8; first = 7
    case ('wed'); np_ = 5; first = 7
    case default
      call die('ERROR: unknown domain type '//trim(w(4)))
    end select
    if (first == 7) read(w(6),*) dom%axis
    do k = 1, np_
      read(w(first+k-1),*) dom%parm(k)
    enddo
  end subroutine read_domain

  ! inDomain (src/domtype.f90:278-434), in single precision as the reference evaluates it
  logical function in_domain(p, dom)
    real(4), intent(in) :: p(3)
    type(domain_t), intent(in) :: dom
    real(4), parameter :: small = tiny(1.0)
    real(4) :: dmin(3), dmax(3), depth, dist, slope, re, ro, u, v
    integer :: a, dir(3), i1, i2
    logical :: good(3)
    in_domain = .false.
    select case (dom%ct)
    case ('rec', 'per')
      dmin = dom%parm(1:5:2); dmax = dom%parm(2:6:2)
      dir = 1
      do a = 1, 3
        if ((abs(dmax(a)) + abs(dmin(a))) < small) dir(a) = 0
        good(a) = dir(a) == 0 .or. (p(a) > dmin(a) .and. p(a) <= dmax(a))
      enddo
      in_domain = all(good)
      if (dom%ct == 'per') then
        depth = dom%parm(7)
        do a = 1, 3
          good(a) = dir(a) == 0 .or. (p(a) > dmin(a) + depth .and. p(a) <= dmax(a) - depth)
        enddo
        if (all(good)) in_domain = .false.
      endif
    case ('sph')
      in_domain = sum((p - dom%parm(1:3))**2.) <= dom%parm(4)**2.
    case ('ell', 'tub')
      call plane(dom%axis, i1, i2)
      u = p(i1) - dom%parm(1); v = p(i2) - dom%parm(2)
      dist = u**2. + v**2.
      slope = (u + v) / (u - v)
      if (abs(u - v) < small) slope = 0.0
      re = dom%parm(3)*dom%parm(4)*sqrt(slope**2.+1.) / sqrt(dom%parm(3)**2.*slope**2.+dom%parm(4)**2.)
      if (dom%ct == 'ell') then
        in_domain = dist <= re**2.
      else
        ro = dom%parm(5)*dom%parm(6)*sqrt(slope**2.+1.) / sqrt(dom%parm(5)**2.*slope**2.+dom%parm(6)**2.)
        in_domain = dist <= re**2. .and. dist > ro**2. .and. &
                    p(dom%axis) > dom%parm(7) .and. p(dom%axis) < dom%parm(8)
      endif
    case ('wed')
      call plane(dom%axis, i1, i2)
      ! the wedge's centre coordinates are parm(i1), parm(i2) (src/domtype.f90:402-421)
      u = p(i1) - dom%parm(i1); v = p(i2) - dom%parm(i2)
      dist = u**2. + v**2.
      if (abs(u - v) < small) then
        re = 0.0
      else
        re = full_angle(u, v)
      endif
      in_domain = dist <= dom%parm(3)**2 .and. re > dom%parm(4) - dom%parm(5) .and. &
                  re <= dom%parm(4)
    end select
    if (.not. dom%inside) in_domain = .not. in_domain
  end function in_domain

  ! the two axes across a domain's axial direction
  subroutine plane(axis, i1, i2)
    integer, intent(in) :: axis
    integer, intent(out) :: i1, i2
    select case (axis)
    case (1); i1 = 2; i2 = 3
    case (2); i1 = 1; i2 = 3
    case default; i1 = 1; i2 = 2
    end select
  end subroutine plane

  ! fullangle (src/domtype.f90:576-598): angle of (u,v) from the first axis in [0, 2 pi)
  real(4) function full_angle(u, v)
    real(4), intent(in) :: u, v
    real(4), parameter :: pi = 3.1415926535, small = tiny(1.0)
    real(4) :: at
    if (abs(u) < small .and. abs(v) < small) then
      at = pi / 2.0
    else
      at = atan2(v, u)
    endif
    full_angle = at
    if (at < 0.0) full_angle = at + 2*pi
  end function full_angle

  ! setType (src/nd.f90:1700-1730): the atoms of a group change type
  subroutine change_type(gname, tname)
    character(*), intent(in) :: gname, tname
    integer :: gid, tid
    gid = group_id(gname)
    tid = type_id(tname)
    if (gid == 0) call die('ERROR: ChangeType: unknown group '//trim(gname))
    if (tid == 0) call die('ERROR: ChangeType: unknown type '//trim(tname))
    where (grp == gid) typ = tid
    state_dirty = .true.
  end subroutine change_type

  ! read_model (src/read_input.f90:277-392): xyz (real coordinates) or LAMMPS dump
  ! (scaled xs in [0,1), single precision as the reference reads them)
  subroutine read_model(fname, kind)
    character(*), intent(in) :: fname, kind
    integer :: u, k, id, ty, e
    real(4) :: lm, lx, x(3)
    character(len=8) :: ctyp
    open(newunit=u, file=fname, action='read', status='old', iostat=e)
    if (e /= 0) call die('read_model: ERROR: opening model file: '//fname)
    write(6,'(A)') '# reading model input file:'//fname//' format: '//kind
    select case (kind)
    case ('xyz', 'XYZ')
      read(u,*) np
      read(u,*)
      call alloc_atoms
      do k = 1, np
        read(u,*,iostat=e) ctyp, pos(1:3,k)
        if (e /= 0) call die('read_model: ERROR: bad model line')
        typ(k) = type_id(ctyp)
        if (typ(k) == 0) call die('ERROR: no atom type defined for atom')
        pos(1:3,k) = pos(1:3,k) / box(1:3)
      enddo
    case ('dump', 'trj')
      read(u,*); read(u,*); read(u,*)
      read(u,*) np
      read(u,*)
      do k = 1, 3
        read(u,*) lm, lx
        box(k) = dble(lx - lm)
      enddo
      read(u,*)
      call alloc_atoms
      do k = 1, np
        read(u,*,iostat=e) id, ty, x
        if (e /= 0) call die('read_model: ERROR: bad model line')
        typ(id) = ty
        pos(1:3,id) = dble(x(1:3) - 0.5)
      enddo
    case default
      call die('ERROR, input Model Type: '//kind//' not a valid input model format')
    end select
    close(u)
    write(6,'(A,i8,A)') '#  read_model: successfully read ', np, ' model nodes.'
  end subroutine read_model

  ! randomPositions (src/nd.f90:934-983): uniform scaled positions in [-0.5,0.5)
  subroutine random_model
    call alloc_atoms
    call random_number(pos)
    pos = pos - 0.5d0
    write(6,'(a,i8)') '#  Number of particles: ', np
  end subroutine random_model

  subroutine set_seed(word)
    character(*), intent(in) :: word
    integer :: n, sd
    integer, allocatable :: seeds(:)
    read(word,*) sd
    call random_seed(size=n)
    allocate(seeds(n))
    seeds = sd
    call random_seed(put=seeds)
  end subroutine set_seed

  ! normalVelocityDistribution (src/nd.f90:1500-1624): v = N(0,1) sqrt(2 kt / (3 m)) per
  ! component (Box-Muller here), then the mean velocity is removed
  ! normalVelocityDistribution (src/nd.f90:1500-1624) with its own normal deviates: the
  ! atoms of group gname (all atoms for 'N/A') get new velocities; the sum of the new
  ! velocities over np is then removed from every atom, as the reference does
  subroutine velocities(kt, gname)
    real(8), intent(in) :: kt
    character(*), intent(in) :: gname
    real(8) :: u1, u2, z(3), vsum(3)
    integer :: k, a, gt
    logical :: ss
    if (.not. have_model) call die('ERROR: Velocity needs a model')
    ss = trim(gname) /= 'N/A'
    gt = 0
    if (ss) gt = group_id(gname)
    if (ss .and. gt == 0) call die('ERROR: Velocity: unknown group '//trim(gname))
    vsum = 0.d0
    do k = 1, np
      if (ss) then
        if (grp(k) /= gt) cycle
      endif
      do a = 1, 3
        call random_number(u1)
        call random_number(u2)
        u1 = max(u1, 1.d-300)
        z(a) = sqrt(-2.d0*log(u1)) * cos(2.d0*3.141592653589793d0*u2)
      enddo
      vel(1:3,k) = z * sqrt(2.d0*kt/(3.d0*mass(typ(k))))
      vsum = vsum + vel(1:3,k)
    enddo
    vsum = vsum / dble(np)
    do a = 1, 3
      vel(a,:) = vel(a,:) - vsum(a)
    enddo
    state_dirty = .true.
  end subroutine velocities

  ! Pairs directive (src/read_input.f90:110-146) and buildPairTables
  subroutine read_pairs
    integer :: k, p1, p2, q, e
    real(8) :: rcut
    logical :: eshift
    integer :: lw
    if (ntypes == 0) call die('ERROR: must first define particle Types before Pairs')
    if (allocated(pairmat)) deallocate(pairmat, pairkind, pairname, pairparams, ptab, ftab, rc2, &
                                       dr2, pairq)
    allocate(pairmat(ntypes,ntypes), pairkind(npairs), pairname(npairs), pairparams(4,npairs), &
             ptab(ptsize+1,npairs), ftab(ptsize+1,npairs), rc2(npairs), dr2(npairs), pairq(npairs))
    pairmat = 0
    pairparams = 0.d0
    pairparams(1,:) = 3.141592653589793d0 / 2.d0
    do k = 1, npairs
      read(5,'(A)',iostat=e) line
      if (e /= 0) call die('#  read_input: ERROR: premature end of input file. No END directive found')
      call strip_comment(line)
      call split(line, w, nw)
      p1 = type_id(w(1))
      p2 = type_id(w(2))
      if (p1 == 0 .or. p2 == 0) call die('ERROR: unknown type in Pairs')
      pairmat(p1,p2) = k
      pairq(k) = charge(p1) * charge(p2)   ! PairChargeProd (src/read_input.f90:134)
      ! 'asym' leaves PairMat(p2,p1) alone; the engine then applies compute1's rules
      ! (src/read_input.f90:130-133, src/nd.f90:155-163)
      if (trim(w(3)) == 'sym') pairmat(p2,p1) = k
      pairname(k) = trim(w(4))
      ! '<name>_shift' (not in the reference): the same pair with its energy table shifted
      ! by V(rc), NDG_ESHIFT; forces are unchanged
      eshift = .false.
      lw = len_trim(w(4))
      if (lw > 6) then
        if (w(4)(lw-5:lw) == '_shift') then
          eshift = .true.
          w(4) = w(4)(1:lw-6)
        endif
      endif
      ! the names pot/dpot accept (src/potentials.f90:81-147): 'morsec', 'buckC' and 'buckc'
      ! take their plain forms there; only MorseC/morseC, BuckC and LJC add the Coulomb term
      select case (trim(w(4)))
      case ('morse', 'Morse', 'morsec')
        pairkind(k) = NDG_MORSE
      case ('MorseC', 'morseC')
        pairkind(k) = NDG_MORSE_C
      case ('BuckC')
        pairkind(k) = NDG_BUCK_C
      case ('buckC', 'buckc')
        pairkind(k) = NDG_BUCK
      case ('LJC')
        pairkind(k) = NDG_LJ_C
      case ('sine')
        pairkind(k) = NDG_SINE
      case ('cosine')
        pairkind(k) = NDG_COSINE
      case ('soft')
        pairkind(k) = NDG_SOFT
      case ('John')
        pairkind(k) = NDG_JOHNSON
      case ('morse_phys')
        pairkind(k) = NDG_MORSE_PHYS
      case ('buck', 'Buck')
        pairkind(k) = NDG_BUCK
      case ('LJ')
        pairkind(k) = NDG_LJ
      case ('spring')
        pairkind(k) = NDG_SPRING
      case default
        call die('ERROR: No potential available for Potential Type: '//trim(w(4)))
      end select
      if (eshift) pairkind(k) = ior(pairkind(k), NDG_ESHIFT)
      do q = 5, nw
        read(w(q),*) pairparams(q-4,k)
      enddo
    enddo
    write(6,'(A)') '# pair potential type matrix'
    rcut = 0.d0
    do k = 1, npairs
      call check(ndg_pair_table_q(int(pairkind(k), c_int), pairparams(:,k), pairq(k), ptrmin, &
                                  int(ptsize, c_int), ptab(:,k), ftab(:,k), rc2(k), dr2(k)))
      write(6,'(a,i3,2f12.6)') '# pair,rcut2,dr2', k, rc2(k), dr2(k)
      rcut = max(rcut, pairparams(1,k))
    enddo
    write(6,'(A)') '# Built potential lists'
    write(6,'(A,g11.4)') '# Global maximum cutoff radius: ', rcut
    if (potouts) call write_tables
    tables_dirty = .true.
    call flush(6)
  end subroutine read_pairs

  subroutine write_tables
    integer :: u1, u2, k
    open(newunit=u1, file='pottab.dat')
    open(newunit=u2, file='forcetab.dat')
    do k = 1, ptsize
      write(u1,*) sqrt(real(k)/real(ptsize)), ptab(k,1:npairs)
      write(u2,*) sqrt(real(k)/real(ptsize)), ftab(k,1:npairs)
    enddo
    close(u1)
    close(u2)
  end subroutine write_tables

  ! write_model (src/read_input.f90:394-431)
  subroutine write_model(u)
    integer, intent(in) :: u
    integer :: k
    select case (trim(outtype))
    case ('XYZ', 'xyz')
      write(u,'(i8)') np
      write(u,'(A)')
      do k = 1, np
        write(u,'(A4,x,7EN14.4)') typename(typ(k)), pos(1:3,k), vel(1:3,k), pene(k)
      enddo
    case ('dump', 'trj')
      write(u,'(A)') 'ITEM: TIMESTEP'
      write(u,'(i7)') step
      write(u,'(A)') 'ITEM: NUMBER OF ATOMS'
      write(u,'(i7)') np
      write(u,'(A)') 'ITEM: BOX BOUNDS pp pp pp'
      write(u,'(2f12.4)') 0.0, box(1)
      write(u,'(2f12.4)') 0.0, box(2)
      write(u,'(2f12.4)') 0.0, box(3)
      write(u,'(A)') 'ITEM: ATOMS id type xs ys zs'
      do k = 1, np
        write(u,'(i7,x,i2,3f9.5)') k, typ(k), real(pos(1:3,k)) + 0.0
      enddo
    case default
      call die('ERROR: model out type: '//trim(outtype)//' not supported')
    end select
    call flush(u)
  end subroutine write_model

  subroutine pull_state
    call check(ndg_get_state(h, pos, vel, acc, frc, pene, box))
  end subroutine pull_state

  ! dynamics (src/nd.f90:60-199) on the device
  subroutine dynamics
    type(ndg_run_params) :: rp
    type(ndg_thermo), allocatable :: rec(:)
    integer(c_int) :: nrec, nup, stepdev, stepsnap, fmt
    integer :: u, left, chunk, k, cur, mcell(3), ne
    integer(8) :: c0, c1, crate
    logical :: frame
    real(8) :: rcut, sk
    character(len=:), allocatable :: fname, names
    if (.not. have_model) call random_model
    if (npairs == 0) call die('ERROR: no Pairs defined')
    if (.not. c_associated(h)) then
      call check(ndg_create(h, int(np, c_int), int(ntypes, c_int), int(npairs, c_int), &
                            int(ptsize, c_int), int(neimax, c_int), int(path, c_int)))
      tables_dirty = .true.
      state_dirty = .true.
    endif
    if (tables_dirty) then
      call check(ndg_set_tables(h, ptab, ftab, rc2, dr2, ptrmin, pairmat, pairparams, mass))
      call check(ndg_set_pair_kinds(h, pairkind))
      tables_dirty = .false.
    endif
    if (state_dirty) then
      call check(ndg_set_state(h, pos, vel, acc, typ, box))
      state_dirty = .false.
    endif
    if (initpos_pending) then
      call check(ndg_init_positions(h))
      initpos_pending = .false.
    endif
    if (holds_dirty) then
      call alloc_groups(.false.)
      call check(ndg_set_holds(h, hold, int(nhold, c_int), bcdisp, bcdir, bcparm))
      holds_dirty = .false.
    endif
    rcut = maxval(pairparams(1,:))
    sk = skin
    if (sk < 0.d0) sk = rcut * 0.2d0
    mcell = int(box / (rcut + sk))
    ne = neimax
    if (ne == 0) ne = int(4./3.*3.141592653589793d0*((rcut+sk)*1.1)**3*np/(box(1)*box(2)*box(3)))
    write(6,'(A,3i4)') '# Number of link cell: ', mcell
    write(6,'(A,i4)') '# Estimating max number of neighbours: ', ne
    if (confouts == 0) confouts = step_num
    if (trim(outtype) == 'NA') outtype = 'xyz'
    write(6,'(a)') ''
    write(6,'(a)') '#  Computing initial forces and energies.'
    write(6,'(a,i8)') '#  Number of time steps: ', step_num
    write(6,'(a,g14.6)') '#  Time step increment: ', dt
    write(6,'(a,g14.6)') '#  Cutoff radius: ', rcut
    write(6,'(a,g14.6)') '#  Cutoff skin: ', sk
    write(6,'(a)') '# Model Output format: '//trim(outtype)
    write(6,'(a)') ''
    write(6,'(a)') '#  Step   Potential     Kinetic        EnergyError    pressure       volume'
    call flush(6)
    rp%nsteps = step_num
    rp%dt = dt
    rp%out_steps = outsteps
    rp%skin = skin
    rp%renei = merge(1, 0, renei)
    rp%renei_freq = renei_freq
    rp%trigger = trigger
    rp%isokinetic = merge(1, 0, isokin); rp%ke_target = ketarget; rp%ke_tau = keparm
    rp%ke_freq = kefreq
    rp%isobaric = merge(1, 0, isobar); rp%p_target = ptarget; rp%p_tau = pparm; rp%p_freq = pfreq
    rp%isoenergetic = merge(1, 0, isoen); rp%e_tau = eparm; rp%e_freq = efreq
    rp%box_load = merge(1, 0, boxload); rp%box_vel = boxvel
    call system_clock(c0, crate)
    select case (trim(outtype))
    case ('XYZ', 'xyz')
      fmt = NDG_FRAME_XYZ
    case ('dump', 'trj')
      fmt = NDG_FRAME_DUMP
    case default
      call die('ERROR: model out type: '//trim(outtype)//' not supported')
    end select
    fname = 'out.'//trim(outtype)
    names = ''
    do k = 1, ntypes
      names = names//typename(k)
    enddo
    if (.not. engine_frames) open(newunit=u, file=fname, position='append')
    call check(ndg_run_begin(h, rp, int(step, c_int)))
    ! the frame before the time loop (src/nd.f90:126-127)
    if (engine_frames) then
      call check(ndg_frame_write(h, fname//c_null_char, fmt, names, int(len(typename), c_int)))
    else
      call check(ndg_snapshot(h))
      call check(ndg_snapshot_get(h, pos, vel, frc, pene, box, stepsnap))
      call write_model(u)
    endif
    left = step_num
    cur = step
    ! Steps run in chunks that end at the ConfOutFreq frames.  A frame is copied on the
    ! device behind its chunk, and the next chunk is queued before the frame is formatted,
    ! so the device keeps stepping while the host writes (src/nd.f90:169-172).
    chunk = min(left, confouts - mod(cur, confouts))
    call check(ndg_run_enqueue(h, int(chunk, c_int)))
    do while (left > 0)
      frame = mod(cur + chunk, confouts) == 0
      if (frame .and. engine_frames) then
        call check(ndg_frame_write(h, fname//c_null_char, fmt, names, int(len(typename), c_int)))
      else if (frame) then
        call check(ndg_snapshot(h))
      endif
      allocate(rec(chunk / max(1, outsteps) + 2))
      call check(ndg_run_wait(h, rec, int(size(rec), c_int), nrec))
      do k = 1, nrec
        write(6, '(x,i8,x,g14.6,x,g14.6,x,g14.6,x,g14.6,x,g14.6)') rec(k)%step, rec(k)%pe, &
          rec(k)%ke, rec(k)%eerr, rec(k)%pressure, rec(k)%volume
      enddo
      call flush(6)
      deallocate(rec)
      cur = cur + chunk
      left = left - chunk
      step = cur
      if (left > 0) then
        chunk = min(left, confouts - mod(cur, confouts))
        call check(ndg_run_enqueue(h, int(chunk, c_int)))
      endif
      if (frame .and. .not. engine_frames) then
        call check(ndg_snapshot_get(h, pos, vel, frc, pene, box, stepsnap))
        call write_model(u)
      endif
    enddo
    call check(ndg_run_end(h, nup, stepdev))
    step = stepdev
    call pull_state
    if (engine_frames) then
      call check(ndg_frame_flush(h))
    else
      close(u)
    endif
    call system_clock(c1)
    write(6, '(a,i8)') '# Number of neighbour list updates:', nup
    write(6, '(a)') ''
    write(6, '(a)') '#  Elapsed time for main dynamics:'
    write(6, '(a,g14.6,a)') '#  ', dble(c1 - c0) / dble(crate), ' seconds'
    call flush(6)
  end subroutine dynamics

  subroutine help
    write(6,'(A)') 'Directives (reference grammar): BoxSize Model Seed NTotal Velocity ReNeighbor'
    write(6,'(A)') ' Skin NeiMax Types Pairs WritePotentialTable Isokinetic Anisokinetic'
    write(6,'(A)') ' Isoenergetic Anisoenergetic Isobaric Anisobaric BoxVel StopBoxVel Exec Print'
    write(6,'(A)') ' OutFreq StepNumber ConfOutFreq Run END; engine: Precision mixed|fp64,'
    write(6,'(A)') ' ListTrigger compat|skin, FrameWriter engine|fortran; Group/ClearGroups/Hold/'
    write(6,'(A)') ' ClearHolds/Disp/ChangeType/InitPos/GroupBufferSize/HoldBufferSize.'
    write(6,'(A)') ' Pair kinds: morse morsec MorseC morse_phys buck BuckC LJ LJC spring sine cosine'
    write(6,'(A)') ' soft John (sym or asym).'
    call flush(6)
  end subroutine help

end program ndyn_gpu
