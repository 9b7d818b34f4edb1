! Drop-in replacement for the reference's time loop, `subroutine driver`
! (src/driver.f90:3-36), that runs the explicit time step on the GPU through the
! eqdyna_amd C ABI (include/eqdyna_amd.h).
!
! Link this file instead of src/driver.f90 and the routines only the time loop
! uses (src/assembleGlobalKU.f90, calcElemKU.f90, calcB.f90, calcElemMass.f90,
! qconstant.f90, hrglss.f90, faulting.f90, fric.f90, comdampv.f90,
! thermop.f90), plus libeqd.so.  Everything else -- input readers, meshgen,
! assembleGlobalMass, init_vel and the output writers -- stays the reference's.
! On return the module globalvar holds what the reference's own driver leaves
! there (fric, fnft, velArr, dispArr, v1, nodalForceArr, stressArr, pstrain,
! the station histories, nt and timeElapsed), so output_frt and the station
! writers run unchanged.
!
! One MPI rank per GPU; several ranks would need the halo exchange of
! MPI4NodalQuant inside the step, which this build does not do.
module eqd_abi
    use iso_c_binding
    implicit none

    type, bind(C) :: eqd_scalars
        real(c_double) :: dt, rdampk, rdampm, w, grav, roumax, rhow, gamar
        real(c_double) :: ccosphi, sinphi, tv, slipRateThres
        real(c_double) :: xsource, ysource, zsource, nucR, nucRuptVel, nucdtau0, nucT
        real(c_double) :: kapa_hg, dx, R, vmaxPML, PMLb(8), pi
        integer(c_int32_t) :: nstep, nn, ne, neq, size_eqidx, size_stress, nf, non, noff, nPML
        integer(c_int32_t) :: C_elastic, C_Q, C_hg, C_nuclea, friclaw, TPV, insertFaultType, nucfault
        integer(c_int32_t) :: precision, device
        real(c_double) :: fric_tp_h
    end type eqd_scalars

    ! field ids of enum eqd_field
    integer(c_int32_t), parameter :: EQD_MESHCOOR = 1, EQD_EQNUMINDEXARR = 2, EQD_EQNUMSTARTINDEXLOC = 3, &
        EQD_NUMOFDOFPERNODEARR = 4, EQD_NODEELEMIDRELATION = 5, EQD_ELEMTYPEARR = 6, EQD_MAT = 7, &
        EQD_ELESHP = 8, EQD_ELEDET = 9, EQD_ELEMASS = 10, EQD_SS = 11, EQD_PHI = 12, &
        EQD_STRESSCOMPINDEXARR = 13, EQD_ELEPOREP = 14, EQD_NODALMASSARR = 15, EQD_FNMS = 16, &
        EQD_NSMP = 17, EQD_UN = 18, EQD_US = 19, EQD_UD = 20, EQD_ARN = 21, EQD_ANONFS = 22, &
        EQD_IDHIST = 23, EQD_STRESSARR = 24, EQD_PSTRAIN = 25, EQD_V1 = 26, EQD_VELARR = 27, &
        EQD_DISPARR = 28, EQD_NODALFORCEARR = 29, EQD_FRIC = 30, EQD_FNFT = 31, EQD_ONFAULTHIST = 32, &
        EQD_OFFFAULTHIST = 33
    integer(c_int32_t), parameter :: EQD_HYP_VALS = 19

    interface
        integer(c_int) function eqd_create(s, h) bind(C, name='eqd_create')
            import :: eqd_scalars, c_ptr, c_int
            type(eqd_scalars), intent(in) :: s
            type(c_ptr), intent(out) :: h
        end function eqd_create
        integer(c_int) function eqd_set(h, field, host, count) bind(C, name='eqd_set')
            import :: c_ptr, c_int, c_int32_t, c_int64_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: field
            type(*), dimension(*), intent(in) :: host
            integer(c_int64_t), value :: count
        end function eqd_set
        integer(c_int) function eqd_commit(h) bind(C, name='eqd_commit')
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function eqd_commit
        integer(c_int) function eqd_run(h, nsteps) bind(C, name='eqd_run')
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nsteps
        end function eqd_run
        integer(c_int) function eqd_run_async(h, nsteps) bind(C, name='eqd_run_async')
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nsteps
        end function eqd_run_async
        integer(c_int) function eqd_wait(h) bind(C, name='eqd_wait')
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function eqd_wait
        integer(c_int) function eqd_get(h, field, host, count) bind(C, name='eqd_get')
            import :: c_ptr, c_int, c_int32_t, c_int64_t
            type(c_ptr), value :: h
            integer(c_int32_t), value :: field
            type(*), dimension(*) :: host
            integer(c_int64_t), value :: count
        end function eqd_get
        integer(c_int) function eqd_record(h, node_ids, n_nodes, every, capacity) bind(C, name='eqd_record')
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), dimension(*), intent(in) :: node_ids
            integer(c_int32_t), value :: n_nodes, every, capacity
        end function eqd_record
        integer(c_int) function eqd_record_get(h, vel, slip_rate, max_records, cnt) bind(C, name='eqd_record_get')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            real(c_double), dimension(*) :: vel, slip_rate
            integer(c_int32_t), value :: max_records
            integer(c_int32_t), intent(out) :: cnt
        end function eqd_record_get
        integer(c_int) function eqd_hypocentre_pairs(h, pairs, cap, n) bind(C, name='eqd_hypocentre_pairs')
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), dimension(*) :: pairs
            integer(c_int32_t), value :: cap
            integer(c_int32_t), intent(out) :: n
        end function eqd_hypocentre_pairs
        integer(c_int) function eqd_hypocentre_get(h, first_step, cnt, out) bind(C, name='eqd_hypocentre_get')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: first_step, cnt
            real(c_double), dimension(*) :: out
        end function eqd_hypocentre_get
        integer(c_int) function eqd_clock(h, nt, t) bind(C, name='eqd_clock')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: nt
            real(c_double), intent(out) :: t
        end function eqd_clock
        integer(c_int) function eqd_nan_info(h, node, nt) bind(C, name='eqd_nan_info')
            import :: c_ptr, c_int, c_int32_t
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: node, nt
        end function eqd_nan_info
        type(c_ptr) function eqd_last_error(h) bind(C, name='eqd_last_error')
            import :: c_ptr
            type(c_ptr), value :: h
        end function eqd_last_error
        integer(c_int) function eqd_destroy(h) bind(C, name='eqd_destroy')
            import :: c_ptr, c_int
            type(c_ptr), value :: h
        end function eqd_destroy
    end interface

contains

    subroutine eqd_check(h, rc, what)
        type(c_ptr), intent(in) :: h
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: i, n
        type(c_ptr) :: p
        if (rc == 0) return
        p = eqd_last_error(h)
        call c_f_pointer(p, msg, [1024])
        n = 0
        do i = 1, 1024
            if (msg(i) == c_null_char) exit
            n = i
        enddo
        write(*,*) 'eqdyna_amd: ', what, ' failed (', rc, '): ', (msg(i), i = 1, n)
        stop 4001
    end subroutine eqd_check

end module eqd_abi

subroutine driver
    use globalvar
    use eqd_abi
    implicit none
    type(eqd_scalars) :: s
    type(c_ptr) :: h
    integer(c_int) :: rc
    integer(c_int32_t) :: ntc, badnode, badnt, chunk, done, nrec, irec, inode, nhyp
    integer(c_int32_t), parameter :: maxrec = 32  ! the records of two chunks
    real(c_double) :: tc
    real(c_double), allocatable :: recv(:, :, :), recs(:, :), hypv(:, :, :)
    integer(c_int32_t), allocatable :: hypp(:)
    integer(c_int32_t) :: nopairs(1)
    integer(c_int64_t) :: nf, non, noff
    character(len=8) :: envval

    nf = nftnd(1)
    non = min(numOfOnFaultStCount, size(anonfs, 2))
    noff = numOfOffFaultStCount
    s%dt = dt; s%rdampk = rdampk; s%rdampm = rdampm; s%w = w; s%grav = grav
    s%roumax = roumax; s%rhow = rhow; s%gamar = gamar; s%ccosphi = ccosphi; s%sinphi = sinphi
    s%tv = tv; s%slipRateThres = slipRateThres; s%xsource = xsource; s%ysource = ysource
    s%zsource = zsource; s%nucR = nucR; s%nucRuptVel = nucRuptVel; s%nucdtau0 = nucdtau0
    s%nucT = nucT; s%kapa_hg = kapa_hg; s%dx = dx; s%R = R; s%vmaxPML = vmaxPML
    s%PMLb = PMLb; s%pi = pi
    s%nstep = nstep; s%nn = totalNumOfNodes; s%ne = totalNumOfElements; s%neq = totalNumOfEquations
    s%size_eqidx = sizeOfEqNumIndexArr; s%size_stress = sizeOfStressDofIndexArr
    s%nf = nftnd(1); s%non = numOfOnFaultStCount; s%noff = numOfOffFaultStCount; s%nPML = nPML
    s%C_elastic = C_elastic; s%C_Q = C_Q; s%C_hg = C_hg; s%C_nuclea = C_nuclea
    s%friclaw = friclaw; s%TPV = TPV; s%insertFaultType = insertFaultType; s%nucfault = nucfault
    s%precision = 0; s%device = 0
    s%fric_tp_h = fric_tp_h
    if (ntotft /= 1) stop 'eqdyna_amd: one fault (ntotft = 1) is supported'
    if (totalNumOfMPIProcs /= 1) stop 'eqdyna_amd: one MPI rank per GPU (npx = npy = npz = 1)'

    rc = eqd_create(s, h)
    if (rc /= 0) stop 'eqdyna_amd: eqd_create rejected the scalars'
    call eqd_check(h, eqd_set(h, EQD_MESHCOOR, meshCoor, int(size(meshCoor), c_int64_t)), 'set meshCoor')
    call eqd_check(h, eqd_set(h, EQD_EQNUMINDEXARR, eqNumIndexArr, int(sizeOfEqNumIndexArr, c_int64_t)), 'set eqNumIndexArr')
    call eqd_check(h, eqd_set(h, EQD_EQNUMSTARTINDEXLOC, eqNumStartIndexLoc, int(totalNumOfNodes, c_int64_t)), 'set eqNumStartIndexLoc')
    call eqd_check(h, eqd_set(h, EQD_NUMOFDOFPERNODEARR, numOfDofPerNodeArr, int(totalNumOfNodes, c_int64_t)), 'set numOfDofPerNodeArr')
    call eqd_check(h, eqd_set(h, EQD_NODEELEMIDRELATION, nodeElemIdRelation, int(size(nodeElemIdRelation), c_int64_t)), 'set IEN')
    call eqd_check(h, eqd_set(h, EQD_ELEMTYPEARR, elemTypeArr, int(totalNumOfElements, c_int64_t)), 'set elemTypeArr')
    call eqd_check(h, eqd_set(h, EQD_MAT, mat, int(size(mat), c_int64_t)), 'set mat')
    call eqd_check(h, eqd_set(h, EQD_ELESHP, eleshp, int(size(eleshp), c_int64_t)), 'set eleshp')
    call eqd_check(h, eqd_set(h, EQD_ELEDET, eledet, int(totalNumOfElements, c_int64_t)), 'set eledet')
    call eqd_check(h, eqd_set(h, EQD_ELEMASS, elemass, int(size(elemass), c_int64_t)), 'set elemass')
    call eqd_check(h, eqd_set(h, EQD_SS, ss, int(size(ss), c_int64_t)), 'set ss')
    call eqd_check(h, eqd_set(h, EQD_PHI, phi, int(size(phi), c_int64_t)), 'set phi')
    call eqd_check(h, eqd_set(h, EQD_STRESSCOMPINDEXARR, stressCompIndexArr, int(totalNumOfElements, c_int64_t)), 'set sci')
    call eqd_check(h, eqd_set(h, EQD_ELEPOREP, eleporep, int(totalNumOfElements, c_int64_t)), 'set eleporep')
    call eqd_check(h, eqd_set(h, EQD_NODALMASSARR, nodalMassArr, int(totalNumOfEquations, c_int64_t)), 'set mass')
    call eqd_check(h, eqd_set(h, EQD_FNMS, fnms, int(totalNumOfNodes, c_int64_t)), 'set fnms')
    call eqd_check(h, eqd_set(h, EQD_NSMP, nsmp, 2 * nf), 'set nsmp')
    call eqd_check(h, eqd_set(h, EQD_UN, un, 3 * nf), 'set un')
    call eqd_check(h, eqd_set(h, EQD_US, us, 3 * nf), 'set us')
    call eqd_check(h, eqd_set(h, EQD_UD, ud, 3 * nf), 'set ud')
    call eqd_check(h, eqd_set(h, EQD_ARN, arn, nf), 'set arn')
    if (non > 0) call eqd_check(h, eqd_set(h, EQD_ANONFS, anonfs, 3 * non), 'set anonfs')
    if (noff > 0) call eqd_check(h, eqd_set(h, EQD_IDHIST, idhist, 3 * 6 * noff), 'set idhist')
    call eqd_check(h, eqd_set(h, EQD_STRESSARR, stressArr, int(sizeOfStressDofIndexArr, c_int64_t)), 'set stressArr')
    call eqd_check(h, eqd_set(h, EQD_PSTRAIN, pstrain, int(totalNumOfElements, c_int64_t)), 'set pstrain')
    call eqd_check(h, eqd_set(h, EQD_V1, v1, int(totalNumOfEquations, c_int64_t)), 'set v1')
    call eqd_check(h, eqd_set(h, EQD_VELARR, velArr, int(size(velArr), c_int64_t)), 'set velArr')
    call eqd_check(h, eqd_set(h, EQD_DISPARR, dispArr, int(size(dispArr), c_int64_t)), 'set dispArr')
    call eqd_check(h, eqd_set(h, EQD_NODALFORCEARR, nodalForceArr, int(totalNumOfEquations, c_int64_t)), 'set F')
    call eqd_check(h, eqd_set(h, EQD_FRIC, fric, 100 * nf), 'set fric')
    call eqd_check(h, eqd_set(h, EQD_FNFT, fnft, nf), 'set fnft')
    call eqd_check(h, eqd_commit(h), 'commit')

    ! Ground motion and source evolution (src/driver.f90:30-33): the library
    ! records the surface nodes' velocities and fric(47) on the device at
    ! every step with mod(nt,10) == 1; after each chunk the records are
    ! appended to the files output_gm and output_src_evol write
    ! (src/library_output.f90:267-312: stream binary, per record the three
    ! velocity components of each surface node, and fric(47) of each pair).
    ! Those routines re-open their still-connected units at every call, which
    ! this compiler's runtime rejects, so the files are written here, one OPEN
    ! per chunk, with the same content.
    if (outputGroundMotion == 1) then
        allocate(recv(3, max(surface_nnode, 1), maxrec), recs(max(nftnd(1), 1), maxrec))
        call eqd_check(h, eqd_record(h, surfaceNodeIdArr, int(surface_nnode, c_int32_t), 10_c_int32_t, maxrec), &
                       'record')
    endif
    ! showSourceDynamics (src/faulting.f90:343-365): the library keeps the
    ! hypocentre pairs' values per step; they are printed after each chunk.
    call eqd_check(h, eqd_hypocentre_pairs(h, nopairs, 0_c_int32_t, nhyp), 'hypocentre pairs')
    allocate(hypp(max(nhyp, 1)), hypv(EQD_HYP_VALS, max(nhyp, 1), 100))
    if (nhyp > 0) call eqd_check(h, eqd_hypocentre_pairs(h, hypp, nhyp, nhyp), 'hypocentre pairs')
    ! Advance in chunks that end where the reference prints progress
    ! (src/driver.f90:13-17), at most 100 steps (10 records).  Each chunk goes
    ! to the device before the records of the one before it are written, so
    ! the files are written while the device steps (eqd_run_async).
    done = 0
    chunk = min(nstep, 100_c_int32_t)
    call progress(done)
    call eqd_check(h, eqd_run_async(h, chunk), 'run')
    do while (chunk > 0)
        rc = eqd_wait(h)
        if (rc == -2) then
            rc = eqd_nan_info(h, badnode, badnt)
            write(*,*) 'Velocity NaN at point ', meshCoor(1,badnode), meshCoor(2,badnode), meshCoor(3,badnode), &
                       ' at time step ', badnt
            stop
        endif
        call eqd_check(h, rc, 'run')
        done = done + chunk
        ! before the next chunk's progress lines, where the reference prints them
        if (nhyp > 0) call source_dynamics(done - chunk + 1, chunk)
        rc = eqd_clock(h, ntc, tc)
        nt = ntc
        timeElapsed = tc
        chunk = min(nstep - done, int(100 - mod(done, 100), c_int32_t))
        if (chunk > 0) then
            call progress(done)
            call eqd_check(h, eqd_run_async(h, chunk), 'run')
        endif
        if (outputGroundMotion == 1) then
            ! the records of the chunk that just ended (not of the one in flight)
            call eqd_check(h, eqd_record_get(h, recv, recs, maxrec, nrec), 'record_get')
            if (nrec > 0 .and. surface_nnode > 0) then
                open(unit=10009+me, file='gm'//mm, status='unknown', position='append', access='stream')
                do irec = 1, nrec
                    do inode = 1, surface_nnode
                        write(10009+me) recv(1, inode, irec), recv(2, inode, irec), recv(3, inode, irec)
                    enddo
                enddo
                close(10009+me)
            endif
            if (nrec > 0 .and. nftnd(1) > 0) then
                open(unit=30009+me, file='src_evol'//mm, position='append', access='stream')
                do irec = 1, nrec
                    do inode = 1, nftnd(1)
                        write(30009+me) recs(inode, irec)
                    enddo
                enddo
                close(30009+me)
            endif
        endif
    enddo

    call eqd_check(h, eqd_get(h, EQD_FRIC, fric, 100 * nf), 'get fric')
    call eqd_check(h, eqd_get(h, EQD_FNFT, fnft, nf), 'get fnft')
    call eqd_check(h, eqd_get(h, EQD_VELARR, velArr, int(size(velArr), c_int64_t)), 'get velArr')
    call eqd_check(h, eqd_get(h, EQD_DISPARR, dispArr, int(size(dispArr), c_int64_t)), 'get dispArr')
    call eqd_check(h, eqd_get(h, EQD_V1, v1, int(totalNumOfEquations, c_int64_t)), 'get v1')
    call eqd_check(h, eqd_get(h, EQD_NODALFORCEARR, nodalForceArr, int(totalNumOfEquations, c_int64_t)), 'get F')
    call eqd_check(h, eqd_get(h, EQD_STRESSARR, stressArr, int(sizeOfStressDofIndexArr, c_int64_t)), 'get stressArr')
    call eqd_check(h, eqd_get(h, EQD_PSTRAIN, pstrain, int(totalNumOfElements, c_int64_t)), 'get pstrain')
    call eqd_check(h, eqd_get(h, EQD_ONFAULTHIST, onFaultQuantHistSCECForm, &
                              int(size(onFaultQuantHistSCECForm), c_int64_t)), 'get on-fault stations')
    if (noff > 0) call eqd_check(h, eqd_get(h, EQD_OFFFAULTHIST, OffFaultStGramSCEC, &
                                             int(size(OffFaultStGramSCEC), c_int64_t)), 'get off-fault stations')
    rc = eqd_destroy(h)
    call get_environment_variable('EQD_FULL_PRECISION_OUTPUT', envval)
    if (trim(envval) == '1') call output_full_precision

contains

    ! the progress lines of src/driver.f90:13-17, before the step after `done`
    subroutine progress(done)
        integer(c_int32_t), intent(in) :: done
        if (me == masterProcsId .and. mod(done + 1, 100) == 1) then
            write(*,*) '=     Current time in dynamic rupture (GPU)                         ='
            write(*,'(X,A,40X,f7.3,4X,A)') '=', timeElapsed + dt, 's'
        endif
    end subroutine progress

    ! the 'Hypocenter dynamics' blocks of src/faulting.f90:350-362 for steps
    ! first .. first + cnt - 1, in step order and pair order
    subroutine source_dynamics(first, cnt)
        integer(c_int32_t), intent(in) :: first, cnt
        integer(c_int32_t) :: k, j, n
        call eqd_check(h, eqd_hypocentre_get(h, first, cnt, hypv), 'hypocentre get')
        do k = 1, cnt
            do j = 1, nhyp
                n = nsmp(1, hypp(j), 1)
                write(*,*) 'Hypocenter dynamics'
                write(*,'(X,A,3E15.7)') 'X, Y, Z      (m)   ', meshCoor(1, n), meshCoor(2, n), meshCoor(3, n)
                write(*,'(X,A,E15.7)') 'TimeElapsed  (s)   ', hypv(1, j, k)
                write(*,'(X,A,3E15.7)') 'n,s,d tract  (MPa) ', hypv(2, j, k)/1.d6, hypv(3, j, k)/1.d6, hypv(4, j, k)/1.d6
                write(*,'(X,A,E15.7)') 'state_normal (MPa) ', hypv(5, j, k)/1.d6
                write(*,'(X,A,3E15.7)') 'n,s,d slip   (m)   ', hypv(6, j, k), hypv(7, j, k), hypv(8, j, k)
                write(*,'(X,A,3E15.7)') 's,d, peak sr (m/s) ', hypv(9, j, k), hypv(10, j, k), hypv(11, j, k)
                write(*,'(X,A,E15.7)') 'cummul slip  (m)   ', hypv(12, j, k)
                write(*,'(X,A,3E15.7)') 'sw_fs, sw_fd, sw_D0', hypv(13, j, k), hypv(14, j, k), hypv(15, j, k)
                write(*,'(X,A,2E15.7)') 'rsf_a, rsf_b       ', hypv(16, j, k), hypv(17, j, k)
                write(*,'(X,A,E15.7)') 'rsf_state          ', hypv(18, j, k)
                write(*,'(X,A,E15.7)') 'Nuc add tau0 (MPa) ', hypv(19, j, k)/1.d6
            enddo
        enddo
    end subroutine source_dynamics
end subroutine driver

! Full-precision companions of the reference's text outputs, written when
! EQD_FULL_PRECISION_OUTPUT=1: the same rows and columns as output_frt
! (src/library_output.f90:160-205), output_onfault_st (:16-96) and
! output_offfault_st (:98-158) in ES25.16E3 (17 significant digits, which
! read back to the same doubles) instead of E18.7 / E16.7, in files named
! like the reference's with '.full' before the extension.  The reference's
! own files are written as before.
subroutine output_full_precision
    use globalvar
    implicit none
    integer (kind = 4) :: i, j, k
    character(len=8) :: cst, cdp, cbd
    character(len=*), parameter :: fmt8 = '(8es25.16e3)', fmt11 = '(11es25.16e3)', fmt7 = '(7es25.16e3)'

    if (nftnd(1) > 0) then
        open(unit=40004+me, file='frt.full.txt'//mm, status='unknown')
        do i = 1, nftnd(1)
            write(40004+me, '(22es25.16e3)') (meshCoor(j, nsmp(1,i,1)), j = 1, 3), fnft(i,1), &
                (fric(j,i,1), j = 71, 76), fric(47,i,1), fric(78,i,1), fric(79,i,1), fric(80,i,1), &
                (fric(j,i,1), j = 31, 36), fric(20,i,1), fric(23,i,1)
        enddo
        close(40004+me)
    endif
    do i = 1, numOfOnFaultStCount
        k = anonfs(3,i)
        if (k /= 1) cycle  ! the reference names main-fault stations only
        write(cst,'(i3.3)') nint(xonfs(1,anonfs(2,i),k)/100.d0)
        write(cdp,'(i3.3)') nint(abs(xonfs(2,anonfs(2,i),k))/dsin(fltxyz(2,4,1))/100.d0)
        open(51, file='faultst'//trim(adjustl(cst))//'dp'//trim(adjustl(cdp))//'.full.txt', status='unknown')
        if (friclaw >= 3) then
            write(51,'(a)') '# t h-slip h-slip-rate h-shear-stress v-slip v-slip-rate v-shear-stress n-stress psi' // &
                            ' temperature pressure'
            do j = 1, nstep
                write(51, fmt11) onFaultQuantHistSCECForm(1,j,i), onFaultQuantHistSCECForm(5,j,i), &
                    onFaultQuantHistSCECForm(2,j,i), onFaultQuantHistSCECForm(8,j,i)/1.0d6, &
                    -onFaultQuantHistSCECForm(6,j,i), -onFaultQuantHistSCECForm(3,j,i), &
                    -onFaultQuantHistSCECForm(9,j,i)/1.0d6, -onFaultQuantHistSCECForm(10,j,i)/1.0d6, &
                    onFaultQuantHistSCECForm(4,j,i), onFaultQuantHistSCECForm(12,j,i), &
                    onFaultQuantHistSCECForm(11,j,i)/1.0d6
            enddo
        else
            write(51,'(a)') '# t h-slip h-slip-rate h-shear-stress v-slip v-slip-rate v-shear-stress n-stress'
            do j = 1, nstep
                write(51, fmt8) onFaultQuantHistSCECForm(1,j,i), onFaultQuantHistSCECForm(5,j,i), &
                    onFaultQuantHistSCECForm(2,j,i), onFaultQuantHistSCECForm(8,j,i)/1.0d6, &
                    -onFaultQuantHistSCECForm(6,j,i), -onFaultQuantHistSCECForm(3,j,i), &
                    -onFaultQuantHistSCECForm(9,j,i)/1.0d6, -onFaultQuantHistSCECForm(10,j,i)/1.0d6
            enddo
        endif
        close(51)
    enddo
    do i = 1, numOfOffFaultStCount
        write(cbd,'(i4.3)') int(x4nds(2,OffFaultStNodeIdIndex(1,i))/100.d0)
        write(cst,'(i4.3)') int(x4nds(1,OffFaultStNodeIdIndex(1,i))/100.d0)
        write(cdp,'(i4.3)') int(abs(x4nds(3,OffFaultStNodeIdIndex(1,i)))/100.d0)
        open(51, file='body'//trim(adjustl(cbd))//'st'//trim(adjustl(cst))//'dp'//trim(adjustl(cdp))//'.full.txt', &
             status='unknown')
        write(51,'(a)') '# t h-disp h-vel v-disp v-vel n-disp n-vel'
        do j = 1, nstep
            write(51, fmt7) OffFaultStGramSCEC(1,j), OffFaultStGramSCEC((i-1)*6+2,j), &
                OffFaultStGramSCEC((i-1)*6+3,j), -OffFaultStGramSCEC((i-1)*6+6,j), &
                -OffFaultStGramSCEC((i-1)*6+7,j), OffFaultStGramSCEC((i-1)*6+4,j), OffFaultStGramSCEC((i-1)*6+5,j)
        enddo
        close(51)
    enddo
end subroutine output_full_precision
