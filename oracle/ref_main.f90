! Test infrastructure, not product code.
!
! Harness main program that drives the reference EQdyna 5.3.3 Fortran, compiled
! from its own sources (see oracle/ref.mk), on one case directory:
!   readInputFiles -> meshgen -> assembleGlobalMass -> driver
! and dumps the reference's FP64 state as raw little-endian arrays plus a
! manifest, so that the C restatement (oracle/eqd_oracle.c) and the HIP path can
! be compared with the reference itself on identical inputs.
!
! Differences from the reference main program (src/eqdyna3d.f90:3-89):
!  * On-fault fields are read from a raw FP64 file on_fault_vars_input.bin
!    (24 fields on the (strike, dip) grid, the field order of
!    src/netcdf_io.f90:41-64) instead of netCDF, and mapped into fric exactly as
!    src/netcdf_io.f90:70-107 does.  netCDF-Fortran is not in this image.
!  * Module knobs that the reference sets by initializers (nPML, R, C_Q, C_hg,
!    rdampm, kapa_hg; src/globalvar.f90:16-75) may be overridden from an
!    optional bKnobs.txt ("key value" lines) before mesh generation.
!  * allocInit / allocInitAfterMeshGen / init_vel are restated here because
!    they live in the reference's main-program file (src/eqdyna3d.f90:91-212).
!  * After the time loop it writes frt.txt<rank> with the reference's own
!    output_frt and, with "dump 1", binary dumps of inputs and outputs.
program eqd_ref_main
    use globalvar
    implicit none
    include 'mpif.h'
    integer (kind = 4) :: ierr, dumpBundle
    real (kind = dp) :: tLoop0, tLoop, tMesh, tMass

    call MPI_Init(ierr)
    call mpi_comm_rank(MPI_COMM_WORLD, me, ierr)
    call mpi_comm_size(MPI_COMM_WORLD, totalNumOfMPIProcs, ierr)

    dumpBundle = 0
    call readglobal
    call readmodelgeometry
    allocate(fxmin(ntotft), fxmax(ntotft), fymin(ntotft), fymax(ntotft), fzmin(ntotft), fzmax(ntotft), &
             material(nmat,n2mat))
    allocate(nonfs(ntotft))
    allocate(fltxyz(2,4,ntotft))
    call readfaultgeometry
    call readmaterial
    call readstations1
    allocate(OffFaultStNodeIdIndex(2,totalNumOfOffSt), xonfs(2,maxval(nonfs),ntotft), x4nds(3,totalNumOfOffSt))
    call readstations2
    if (insertFaultType > 0) call read_fault_rough_geometry
    call ref_load_knobs(dumpBundle)
    call warning

    allocate(nftnd(ntotft), localShapeFunc(nrowsh,nen))
    call calcLocalShapeFunc
    call mesh4num
    call ref_alloc_init
    call MPI_Barrier(MPI_COMM_WORLD, ierr)
    tMesh = MPI_WTIME()
    call meshgen
    tMesh = MPI_WTIME() - tMesh
    call ref_read_on_fault_raw
    if (outputGroundMotion == 1 .or. outputFinalSurfDisp == 1) call ref_find_surface_nodes
    call ref_alloc_after_meshgen
    call MPI_Barrier(MPI_COMM_WORLD, ierr)
    tMass = MPI_WTIME()
    call assembleGlobalMass
    tMass = MPI_WTIME() - tMass
    call ref_init_vel
    call MPI_Allreduce(MPI_IN_PLACE, tMesh, 1, MPI_DOUBLE_PRECISION, MPI_MAX, MPI_COMM_WORLD, ierr)
    call MPI_Allreduce(MPI_IN_PLACE, tMass, 1, MPI_DOUBLE_PRECISION, MPI_MAX, MPI_COMM_WORLD, ierr)

    if (dumpBundle == 1) call ref_dump_bundle('bundle_in')

    call MPI_Barrier(MPI_COMM_WORLD, ierr)
    tLoop0 = MPI_WTIME()
    call driver
    tLoop = MPI_WTIME() - tLoop0
    call MPI_Allreduce(MPI_IN_PLACE, tLoop, 1, MPI_DOUBLE_PRECISION, MPI_MAX, MPI_COMM_WORLD, ierr)

    call output_frt
    if (outputFinalSurfDisp == 1) call output_finalSurfDisp
    if (dumpBundle == 1) call ref_dump_outputs('ref_out')
    if (me == 0) then
        write(*,'(A,I12,A,I12,A,I8,A,ES16.8,A,ES16.8,A,ES16.8)') 'EQD_REF nodes ', totalNumOfNodes, &
            ' elems ', totalNumOfElements, ' steps ', nstep, ' t_loop_s ', tLoop, ' t_meshgen_s ', tMesh, &
            ' t_mass_s ', tMass
    endif
    call MPI_Finalize(ierr)
end program eqd_ref_main

subroutine ref_load_knobs(dumpBundle)
    use globalvar
    implicit none
    integer (kind = 4) :: dumpBundle, ios
    logical :: ex
    character(len=64) :: key
    real (kind = dp) :: val
    inquire(file='bKnobs.txt', exist=ex)
    if (.not. ex) return
    open(unit=77, file='bKnobs.txt', status='old', form='formatted')
    do
        read(77, *, iostat=ios) key, val
        if (ios /= 0) exit
        select case (trim(key))
        case ('nPML')
            nPML = nint(val)
        case ('R')
            R = val
        case ('C_Q')
            C_Q = nint(val)
        case ('C_hg')
            C_hg = nint(val)
        case ('rdampm')
            rdampm = val
        case ('kapa_hg')
            kapa_hg = val
        case ('writeCompTime')
            writeCompTime = nint(val)
        case ('dump')
            dumpBundle = nint(val)
        case default
            write(*,*) 'bKnobs.txt: unknown key ', trim(key)
            stop 9001
        end select
    end do
    close(77)
end subroutine ref_load_knobs

! Restates allocInit, src/eqdyna3d.f90:91-151.
subroutine ref_alloc_init
    use globalvar
    implicit none
    write(mm,'(i6)') me
    mm = trim(adjustl(mm))
    allocate(eqNumIndexArr(sizeOfEqNumIndexArr), eqNumStartIndexLoc(totalNumOfNodes), &
            numOfDofPerNodeArr(totalNumOfNodes), meshCoor(ndof,totalNumOfNodes), &
            fnms(totalNumOfNodes), surfaceNodeIdArr(totalNumOfNodes), &
            nodeElemIdRelation(nen,totalNumOfElements), mat(totalNumOfElements,5), &
            elemTypeArr(totalNumOfElements), eleporep(totalNumOfElements), &
            pstrain(totalNumOfElements), eledet(totalNumOfElements), &
            elemass(nee,totalNumOfElements), eleshp(nrowsh-1,nen,totalNumOfElements), &
            ss(6,totalNumOfElements), phi(nen,4,totalNumOfElements))
    meshCoor = 0.0d0
    fnms     = 0.0d0
    eqNumIndexArr = 0
    eqNumStartIndexLoc = 0
    numOfDofPerNodeArr = 0
    surfaceNodeIdArr   = 0
    nodeElemIdRelation = 0
    elemTypeArr = 0
    mat      = 0.0d0
    eleporep = 0.0d0
    pstrain  = 0.0d0
    eledet   = 0.0d0
    elemass  = 0.0d0
    eleshp   = 0.0d0
    ss       = 0.0d0
    phi      = 0.0d0
    nftmx = maxval(nftnd)
    if (nftmx <= 0) nftmx = 1
    nonmx = sum(nonfs)
    allocate(onFaultTPHist(2,nftmx,nstep,ntotft))
    onFaultTPHist = 0.0d0
    allocate(nsmp(2,nftmx,ntotft), fnft(nftmx,ntotft), un(3,nftmx,ntotft), &
             us(3,nftmx,ntotft), ud(3,nftmx,ntotft), fric(100,nftmx,ntotft), &
             arn(nftmx,ntotft), anonfs(3,nonmx), fltgm(nftmx), &
             Tatnode(nftmx,ntotft), patnode(nftmx,ntotft))
    fltgm   = 0
    nsmp    = 0
    fnft    = 99999.0d0
    fric    = 0.0d0
    un      = 0.0d0
    us      = 1000.0d0
    ud      = 0.0d0
    arn     = 0.0d0
    anonfs  = 0
    Tatnode = 0.0d0
    patnode = 0.0d0
    allocate(stressCompIndexArr(totalNumOfElements))
    allocate(stressArr(5*sizeOfEqNumIndexArr))
    stressCompIndexArr = 0
    stressArr = 0.0d0
end subroutine ref_alloc_init

! Restates allocInitAfterMeshGen, src/eqdyna3d.f90:153-189.
subroutine ref_alloc_after_meshgen
    use globalvar
    implicit none
    integer (kind = 4) :: iSt, iDof, dispOrVel, rowCount, nodeId
    if (numOfOnFaultStCount <= 0) numOfOnFaultStCount = 1
    allocate(onFaultQuantHistSCECForm(12,nstep,numOfOnFaultStCount))
    onFaultQuantHistSCECForm = 0.0d0
    allocate(nodalForceArr(totalNumOfEquations), v1(totalNumOfEquations), &
             nodalMassArr(totalNumOfEquations), &
             velArr(ndof,totalNumOfNodes), dispArr(ndof,totalNumOfNodes))
    nodalForceArr = 0.0d0
    nodalMassArr  = 0.0d0
    v1      = 0.0d0
    velArr  = 0.0d0
    dispArr = 0.0d0
    if (numOfOffFaultStCount > 0) then
        allocate(idhist(3,numOfOffFaultStCount*ndof*2), &
                 OffFaultStGramSCEC(numOfOffFaultStCount*ndof*2+1,nstep))
        idhist = 0
        OffFaultStGramSCEC = 0.0d0
        rowCount = 0
        do iSt = 1, numOfOffFaultStCount
            do iDof = 1, ndof
                do dispOrVel = 1, 2
                    rowCount = rowCount + 1
                    nodeId = OffFaultStNodeIdIndex(2,iSt)
                    idhist(1,rowCount) = nodeId
                    idhist(2,rowCount) = iDof
                    idhist(3,rowCount) = dispOrVel
                enddo
            enddo
        enddo
    endif
end subroutine ref_alloc_after_meshgen

! Restates init_vel, src/eqdyna3d.f90:191-212.
subroutine ref_init_vel
    use globalvar
    implicit none
    integer (kind = 4) :: i, ift, tmp
    do ift = 1, ntotft
        do i = 1, nftnd(ift)
            tmp = eqNumStartIndexLoc(nsmp(1,i,ift))
            v1(eqNumIndexArr(tmp+1)) = fric(34,i,ift)
            v1(eqNumIndexArr(tmp+2)) = fric(35,i,ift)
            v1(eqNumIndexArr(tmp+3)) = fric(36,i,ift)
            tmp = eqNumStartIndexLoc(nsmp(2,i,ift))
            v1(eqNumIndexArr(tmp+1)) = fric(31,i,ift)
            v1(eqNumIndexArr(tmp+2)) = fric(32,i,ift)
            v1(eqNumIndexArr(tmp+3)) = fric(33,i,ift)
        enddo
    enddo
end subroutine ref_init_vel

! On-fault fields from on_fault_vars_input.bin: float64, Fortran shape
! (fnx, fnz, 24) with fnx along strike and fnz along dip, fields in the order of
! src/netcdf_io.f90:41-64.  The mapping into fric follows src/netcdf_io.f90:70-107.
subroutine ref_read_on_fault_raw
    use globalvar
    implicit none
    integer (kind = 4) :: fnx, fnz, i, ii, jj, ift
    real (kind = dp), allocatable :: v(:,:,:)
    real (kind = dp) :: xcord, zcord
    logical :: ex
    fnx = nint((fxmax(1) - fxmin(1))/dx) + 1
    fnz = nint((fzmax(1) - fzmin(1))/dz) + 1
    allocate(v(fnx,fnz,24))
    inquire(file='on_fault_vars_input.bin', exist=ex)
    if (.not. ex) then
        write(*,*) 'on_fault_vars_input.bin is missing'
        stop 9002
    endif
    open(unit=78, file='on_fault_vars_input.bin', access='stream', form='unformatted', status='old')
    read(78) v
    close(78)
    do ift = 1, ntotft
        do i = 1, nftnd(ift)
            xcord = meshCoor(1, nsmp(1,i,ift))
            zcord = meshCoor(3, nsmp(1,i,ift))
            ii = nint((xcord - fxmin(ift))/dx) + 1
            jj = nint((zcord - fzmin(ift))/dz) + 1
            fric(1,  i, 1) = v(ii,jj,1)
            fric(2,  i, 1) = v(ii,jj,2)
            fric(3,  i, 1) = v(ii,jj,3)
            fric(9,  i, 1) = v(ii,jj,4)
            fric(10, i, 1) = v(ii,jj,5)
            fric(11, i, 1) = v(ii,jj,6)
            fric(12, i, 1) = v(ii,jj,7)
            fric(13, i, 1) = v(ii,jj,8)
            fric(14, i, 1) = v(ii,jj,9)
            fric(15, i, 1) = v(ii,jj,10)
            fric(16, i, 1) = v(ii,jj,11)
            fric(17, i, 1) = v(ii,jj,12)
            fric(18, i, 1) = v(ii,jj,13)
            fric(19, i, 1) = v(ii,jj,14)
            fric(40, i, 1) = v(ii,jj,15)
            fric(41, i, 1) = v(ii,jj,16)
            fric(42, i, 1) = v(ii,jj,17)
            fric(46, i, 1) = v(ii,jj,18)
            fric(8,  i, 1) = v(ii,jj,19)
            fric(7,  i, 1) = v(ii,jj,20)
            fric(20, i, 1) = v(ii,jj,21)
            fric(47, i, 1) = fric(46, i, 1)
            fric(25, i, 1) = 0.0d0
            fric(26, i, 1) = fric(46, i, 1)
            fric(27, i, 1) = 0.0d0
            fric(5,  i, 1) = v(ii,jj,22)
            fric(4,  i, 1) = v(ii,jj,23)
            fric(49, i, 1) = v(ii,jj,24)
            fric(23, i, 1) = abs(fric(7, i, 1))
        enddo
    enddo
    deallocate(v)
end subroutine ref_read_on_fault_raw

! ---------------------------------------------------------------- dumps
subroutine dump_open_manifest(dir)
    implicit none
    character(len=*) :: dir
    call execute_command_line('mkdir -p '//trim(dir))
    open(unit=81, file=trim(dir)//'/manifest.txt', status='replace', form='formatted')
end subroutine dump_open_manifest

subroutine dump_r8(dir, name, a, n, nd, d1, d2, d3)
    implicit none
    character(len=*) :: dir, name
    integer (kind = 4) :: n, nd, d1, d2, d3
    real (kind = 8) :: a(n)
    open(unit=82, file=trim(dir)//'/'//trim(name)//'.bin', access='stream', form='unformatted', status='replace')
    write(82) a
    close(82)
    write(81,'(A,1X,A,1X,I2,3(1X,I12))') trim(name), 'f8', nd, d1, d2, d3
end subroutine dump_r8

subroutine dump_i4(dir, name, a, n, nd, d1, d2, d3)
    implicit none
    character(len=*) :: dir, name
    integer (kind = 4) :: n, nd, d1, d2, d3
    integer (kind = 4) :: a(n)
    open(unit=82, file=trim(dir)//'/'//trim(name)//'.bin', access='stream', form='unformatted', status='replace')
    write(82) a
    close(82)
    write(81,'(A,1X,A,1X,I2,3(1X,I12))') trim(name), 'i4', nd, d1, d2, d3
end subroutine dump_i4

subroutine dump_scalar(name, v)
    implicit none
    character(len=*) :: name
    real (kind = 8) :: v
    write(83,'(A,1X,ES26.17E3)') trim(name), v
end subroutine dump_scalar

subroutine ref_dump_bundle(dir)
    use globalvar
    implicit none
    character(len=*) :: dir
    integer (kind = 4) :: nn, ne, neq, nf, noff
    nn = totalNumOfNodes
    ne = totalNumOfElements
    neq = totalNumOfEquations
    nf = nftmx
    noff = numOfOffFaultStCount
    call dump_open_manifest(dir)
    open(unit=83, file=trim(dir)//'/scalars.txt', status='replace', form='formatted')
    call dump_scalar('dt', dt)
    call dump_scalar('nstep', dble(nstep))
    call dump_scalar('totalNumOfNodes', dble(nn))
    call dump_scalar('totalNumOfElements', dble(ne))
    call dump_scalar('totalNumOfEquations', dble(neq))
    call dump_scalar('sizeOfEqNumIndexArr', dble(sizeOfEqNumIndexArr))
    call dump_scalar('sizeOfStressDofIndexArr', dble(sizeOfStressDofIndexArr))
    call dump_scalar('nftnd', dble(nftnd(1)))
    call dump_scalar('nftmx', dble(nftmx))
    call dump_scalar('ntotft', dble(ntotft))
    call dump_scalar('C_elastic', dble(C_elastic))
    call dump_scalar('C_Q', dble(C_Q))
    call dump_scalar('C_hg', dble(C_hg))
    call dump_scalar('C_nuclea', dble(C_nuclea))
    call dump_scalar('friclaw', dble(friclaw))
    call dump_scalar('TPV', dble(TPV))
    call dump_scalar('insertFaultType', dble(insertFaultType))
    call dump_scalar('nucfault', dble(nucfault))
    call dump_scalar('nPML', dble(nPML))
    call dump_scalar('R', R)
    call dump_scalar('vmaxPML', vmaxPML)
    call dump_scalar('PMLb1', PMLb(1))
    call dump_scalar('PMLb2', PMLb(2))
    call dump_scalar('PMLb3', PMLb(3))
    call dump_scalar('PMLb4', PMLb(4))
    call dump_scalar('PMLb5', PMLb(5))
    call dump_scalar('PMLb6', PMLb(6))
    call dump_scalar('PMLb7', PMLb(7))
    call dump_scalar('PMLb8', PMLb(8))
    call dump_scalar('rdampk', rdampk)
    call dump_scalar('rdampm', rdampm)
    call dump_scalar('w', w)
    call dump_scalar('grav', grav)
    call dump_scalar('roumax', roumax)
    call dump_scalar('rhow', rhow)
    call dump_scalar('gamar', gamar)
    call dump_scalar('ccosphi', ccosphi)
    call dump_scalar('sinphi', sinphi)
    call dump_scalar('tv', tv)
    call dump_scalar('slipRateThres', slipRateThres)
    call dump_scalar('xsource', xsource)
    call dump_scalar('ysource', ysource)
    call dump_scalar('zsource', zsource)
    call dump_scalar('nucR', nucR)
    call dump_scalar('nucRuptVel', nucRuptVel)
    call dump_scalar('nucdtau0', nucdtau0)
    call dump_scalar('nucT', nucT)
    call dump_scalar('kapa_hg', kapa_hg)
    call dump_scalar('dx', dx)
    call dump_scalar('dy', dy)
    call dump_scalar('dz', dz)
    call dump_scalar('tol', tol)
    call dump_scalar('pi', pi)
    call dump_scalar('numOfOnFaultStCount', dble(numOfOnFaultStCount))
    call dump_scalar('numOfOffFaultStCount', dble(noff))
    call dump_scalar('mode', dble(mode))
    close(83)
    call dump_r8(dir, 'meshCoor', meshCoor, 3*nn, 2, 3, nn, 1)
    call dump_i4(dir, 'eqNumIndexArr', eqNumIndexArr, sizeOfEqNumIndexArr, 1, sizeOfEqNumIndexArr, 1, 1)
    call dump_i4(dir, 'eqNumStartIndexLoc', eqNumStartIndexLoc, nn, 1, nn, 1, 1)
    call dump_i4(dir, 'numOfDofPerNodeArr', numOfDofPerNodeArr, nn, 1, nn, 1, 1)
    call dump_i4(dir, 'nodeElemIdRelation', nodeElemIdRelation, 8*ne, 2, 8, ne, 1)
    call dump_i4(dir, 'elemTypeArr', elemTypeArr, ne, 1, ne, 1, 1)
    call dump_r8(dir, 'mat', mat, 5*ne, 2, ne, 5, 1)
    call dump_r8(dir, 'eleshp', eleshp, 24*ne, 3, 3, 8, ne)
    call dump_r8(dir, 'eledet', eledet, ne, 1, ne, 1, 1)
    call dump_r8(dir, 'elemass', elemass, 24*ne, 2, 24, ne, 1)
    call dump_r8(dir, 'ss', ss, 6*ne, 2, 6, ne, 1)
    call dump_r8(dir, 'phi', phi, 32*ne, 3, 8, 4, ne)
    call dump_i4(dir, 'stressCompIndexArr', stressCompIndexArr, ne, 1, ne, 1, 1)
    call dump_r8(dir, 'stressArr', stressArr, sizeOfStressDofIndexArr, 1, sizeOfStressDofIndexArr, 1, 1)
    call dump_r8(dir, 'eleporep', eleporep, ne, 1, ne, 1, 1)
    call dump_r8(dir, 'pstrain', pstrain, ne, 1, ne, 1, 1)
    call dump_r8(dir, 'nodalMassArr', nodalMassArr, neq, 1, neq, 1, 1)
    call dump_r8(dir, 'fnms', fnms, nn, 1, nn, 1, 1)
    call dump_r8(dir, 'v1', v1, neq, 1, neq, 1, 1)
    call dump_i4(dir, 'nsmp', nsmp, 2*nf, 2, 2, nf, 1)
    call dump_r8(dir, 'un', un, 3*nf, 2, 3, nf, 1)
    call dump_r8(dir, 'us', us, 3*nf, 2, 3, nf, 1)
    call dump_r8(dir, 'ud', ud, 3*nf, 2, 3, nf, 1)
    call dump_r8(dir, 'arn', arn, nf, 1, nf, 1, 1)
    call dump_r8(dir, 'fric', fric, 100*nf, 2, 100, nf, 1)
    call dump_r8(dir, 'fnft', fnft, nf, 1, nf, 1, 1)
    if (nonmx > 0) call dump_i4(dir, 'anonfs', anonfs, 3*nonmx, 2, 3, nonmx, 1)
    if (noff > 0) call dump_i4(dir, 'idhist', idhist, 3*6*noff, 2, 3, 6*noff, 1)
    close(81)
end subroutine ref_dump_bundle

subroutine ref_dump_outputs(dir)
    use globalvar
    implicit none
    character(len=*) :: dir
    integer (kind = 4) :: nn, ne, neq, nf, noff, non
    nn = totalNumOfNodes
    ne = totalNumOfElements
    neq = totalNumOfEquations
    nf = nftmx
    noff = numOfOffFaultStCount
    non = numOfOnFaultStCount
    call dump_open_manifest(dir)
    call dump_r8(dir, 'fric', fric, 100*nf, 2, 100, nf, 1)
    call dump_r8(dir, 'fnft', fnft, nf, 1, nf, 1, 1)
    call dump_r8(dir, 'velArr', velArr, 3*nn, 2, 3, nn, 1)
    call dump_r8(dir, 'dispArr', dispArr, 3*nn, 2, 3, nn, 1)
    call dump_r8(dir, 'v1', v1, neq, 1, neq, 1, 1)
    call dump_r8(dir, 'nodalForceArr', nodalForceArr, neq, 1, neq, 1, 1)
    call dump_r8(dir, 'stressArr', stressArr, sizeOfStressDofIndexArr, 1, sizeOfStressDofIndexArr, 1, 1)
    call dump_r8(dir, 'pstrain', pstrain, ne, 1, ne, 1, 1)
    call dump_r8(dir, 'onFaultQuantHistSCECForm', onFaultQuantHistSCECForm, 12*nstep*non, 3, 12, nstep, non)
    if (noff > 0) call dump_r8(dir, 'OffFaultStGramSCEC', OffFaultStGramSCEC, (6*noff+1)*nstep, 2, 6*noff+1, nstep, 1)
    close(81)
end subroutine ref_dump_outputs

! find_surfaceNodeIdArr (src/library_output.f90:247-264): the surface nodes
! within 20 km of the fault's extent, in node order, and their coordinates in
! surface_coor.txt<rank>.  Restated because the reference's routine re-opens
! its output unit, still connected, for every node, which this compiler's
! runtime rejects; the selection and the file are the same.
subroutine ref_find_surface_nodes
    use globalvar
    implicit none
    integer (kind = 4) :: i
    logical :: opened
    opened = .false.
    do i = 1, totalNumOfNodes
        if ((meshCoor(1,i) < fltxyz(2,1,1) + 20.0d3) .and. (meshCoor(1,i) > fltxyz(1,1,1) - 20.0d3) &
                .and. (meshCoor(2,i) < fltxyz(2,2,1) + 20.0d3) .and. (meshCoor(2,i) > fltxyz(1,2,1) - 20.0d3) &
                .and. (abs(meshCoor(3,i)) < dx/1000)) then
            surface_nnode = surface_nnode + 1
            surfaceNodeIdArr(surface_nnode) = i
            if (.not. opened) then
                open(unit=10008+me, file='surface_coor.txt'//mm, status='unknown', position='append')
                opened = .true.
            endif
            write(10008+me,'(1x,3e18.7e4)') meshCoor(1,i), meshCoor(2,i), meshCoor(3,i)
        endif
    enddo
    if (opened) close(10008+me)
end subroutine ref_find_surface_nodes
