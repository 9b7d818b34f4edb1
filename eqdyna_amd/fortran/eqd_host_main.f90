! Host program for the GPU time step: the reference's start-up sequence
! (src/eqdyna3d.f90:10-86) with `call driver` resolved to eqd_driver.f90.
!
! Where netCDF-Fortran is installed the reference's own src/eqdyna3d.f90 and
! src/netcdf_io.f90 link against eqd_driver.f90 unchanged.  This image has no
! netCDF, so this program reads the 24 on-fault fields from
! on_fault_vars_input.bin instead (float64, Fortran shape (nfx, nfz, 24), the
! field order of src/netcdf_io.f90:41-64) and maps them into fric as
! src/netcdf_io.f90:70-107 does.  allocInit, allocInitAfterMeshGen and init_vel
! live in the reference's main-program file and are restated here
! (src/eqdyna3d.f90:91-212).
program eqdyna_gpu
    use globalvar
    implicit none
    include 'mpif.h'
    integer (kind = 4) :: ierr
    real (kind = dp) :: t0

    call MPI_Init(ierr)
    call mpi_comm_rank(MPI_COMM_WORLD, me, ierr)
    call mpi_comm_size(MPI_COMM_WORLD, totalNumOfMPIProcs, ierr)
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
    call warning
    allocate(nftnd(ntotft), localShapeFunc(nrowsh,nen))
    call calcLocalShapeFunc
    call mesh4num
    call host_alloc_init
    call meshgen
    call checkMeshMaterial
    call host_read_on_fault
    if (outputGroundMotion == 1 .or. outputFinalSurfDisp == 1) call host_find_surface_nodes
    call host_alloc_after_meshgen
    call assembleGlobalMass
    call host_init_vel
    t0 = MPI_WTIME()
    call driver
    if (me == 0) write(*,'(A,ES16.8)') 'EQD_GPU t_driver_s ', MPI_WTIME() - t0
    call output_onfault_st
    call output_offfault_st
    call output_frt
    if (output_plastic == 1) call output_plastic_strain
    if (outputFinalSurfDisp == 1) call output_finalSurfDisp
    call MPI_Finalize(ierr)
end program eqdyna_gpu

subroutine checkMeshMaterial
    use globalvar
    implicit none
    integer (kind = 4) :: i
    do i = 1, totalNumOfElements
        if (mat(i,1) == 0.0d3 .or. mat(i,2) == 0.0d3 .or. mat(i,3) == 0.0d3) then
            write(*,*) 'Element ', i, ' is not assigned material property. Exiting ... ...'
            stop
        endif
    enddo
end subroutine checkMeshMaterial

subroutine host_alloc_init
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
    meshCoor = 0.0d0; fnms = 0.0d0; eqNumIndexArr = 0; eqNumStartIndexLoc = 0
    numOfDofPerNodeArr = 0; surfaceNodeIdArr = 0; nodeElemIdRelation = 0; elemTypeArr = 0
    mat = 0.0d0; eleporep = 0.0d0; pstrain = 0.0d0; eledet = 0.0d0; elemass = 0.0d0
    eleshp = 0.0d0; ss = 0.0d0; phi = 0.0d0
    nftmx = maxval(nftnd)
    if (nftmx <= 0) nftmx = 1
    nonmx = sum(nonfs)
    allocate(onFaultTPHist(2,nftmx,1,ntotft))  ! thermal pressurization history: not used by the GPU step
    onFaultTPHist = 0.0d0
    allocate(nsmp(2,nftmx,ntotft), fnft(nftmx,ntotft), un(3,nftmx,ntotft), &
             us(3,nftmx,ntotft), ud(3,nftmx,ntotft), fric(100,nftmx,ntotft), &
             arn(nftmx,ntotft), anonfs(3,nonmx), fltgm(nftmx), &
             Tatnode(nftmx,ntotft), patnode(nftmx,ntotft))
    fltgm = 0; nsmp = 0; fnft = 99999.0d0; fric = 0.0d0; un = 0.0d0; us = 1000.0d0; ud = 0.0d0
    arn = 0.0d0; anonfs = 0; Tatnode = 0.0d0; patnode = 0.0d0
    allocate(stressCompIndexArr(totalNumOfElements))
    allocate(stressArr(5*sizeOfEqNumIndexArr))
    stressCompIndexArr = 0
    stressArr = 0.0d0
end subroutine host_alloc_init

subroutine host_alloc_after_meshgen
    use globalvar
    implicit none
    integer (kind = 4) :: iSt, iDof, dispOrVel, rowCount
    if (numOfOnFaultStCount <= 0) numOfOnFaultStCount = 1
    allocate(onFaultQuantHistSCECForm(12,nstep,numOfOnFaultStCount))
    onFaultQuantHistSCECForm = 0.0d0
    allocate(nodalForceArr(totalNumOfEquations), v1(totalNumOfEquations), nodalMassArr(totalNumOfEquations), &
             velArr(ndof,totalNumOfNodes), dispArr(ndof,totalNumOfNodes))
    nodalForceArr = 0.0d0; nodalMassArr = 0.0d0; v1 = 0.0d0; velArr = 0.0d0; dispArr = 0.0d0
    if (numOfOffFaultStCount > 0) then
        allocate(idhist(3,numOfOffFaultStCount*ndof*2), OffFaultStGramSCEC(numOfOffFaultStCount*ndof*2+1,nstep))
        idhist = 0
        OffFaultStGramSCEC = 0.0d0
        rowCount = 0
        do iSt = 1, numOfOffFaultStCount
            do iDof = 1, ndof
                do dispOrVel = 1, 2
                    rowCount = rowCount + 1
                    idhist(1,rowCount) = OffFaultStNodeIdIndex(2,iSt)
                    idhist(2,rowCount) = iDof
                    idhist(3,rowCount) = dispOrVel
                enddo
            enddo
        enddo
    endif
end subroutine host_alloc_after_meshgen

subroutine host_init_vel
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
end subroutine host_init_vel

subroutine host_read_on_fault
    use globalvar
    implicit none
    integer (kind = 4) :: fnx, fnz, i, ii, jj, ift
    real (kind = dp), allocatable :: v(:,:,:)
    logical :: ex
    fnx = nint((fxmax(1) - fxmin(1))/dx) + 1
    fnz = nint((fzmax(1) - fzmin(1))/dz) + 1
    allocate(v(fnx,fnz,24))
    inquire(file='on_fault_vars_input.bin', exist=ex)
    if (.not. ex) stop 'on_fault_vars_input.bin is missing'
    open(unit=78, file='on_fault_vars_input.bin', access='stream', form='unformatted', status='old')
    read(78) v
    close(78)
    do ift = 1, ntotft
        do i = 1, nftnd(ift)
            ii = nint((meshCoor(1, nsmp(1,i,ift)) - fxmin(ift))/dx) + 1
            jj = nint((meshCoor(3, nsmp(1,i,ift)) - fzmin(ift))/dz) + 1
            fric(1,i,1) = v(ii,jj,1);   fric(2,i,1) = v(ii,jj,2);   fric(3,i,1) = v(ii,jj,3)
            fric(9,i,1) = v(ii,jj,4);   fric(10,i,1) = v(ii,jj,5);  fric(11,i,1) = v(ii,jj,6)
            fric(12,i,1) = v(ii,jj,7);  fric(13,i,1) = v(ii,jj,8);  fric(14,i,1) = v(ii,jj,9)
            fric(15,i,1) = v(ii,jj,10); fric(16,i,1) = v(ii,jj,11); fric(17,i,1) = v(ii,jj,12)
            fric(18,i,1) = v(ii,jj,13); fric(19,i,1) = v(ii,jj,14); fric(40,i,1) = v(ii,jj,15)
            fric(41,i,1) = v(ii,jj,16); fric(42,i,1) = v(ii,jj,17); fric(46,i,1) = v(ii,jj,18)
            fric(8,i,1) = v(ii,jj,19);  fric(7,i,1) = v(ii,jj,20);  fric(20,i,1) = v(ii,jj,21)
            fric(47,i,1) = fric(46,i,1)
            fric(25,i,1) = 0.0d0
            fric(26,i,1) = fric(46,i,1)
            fric(27,i,1) = 0.0d0
            fric(5,i,1) = v(ii,jj,22);  fric(4,i,1) = v(ii,jj,23);  fric(49,i,1) = v(ii,jj,24)
            fric(23,i,1) = abs(fric(7,i,1))
        enddo
    enddo
    deallocate(v)
end subroutine host_read_on_fault

! find_surfaceNodeIdArr (src/library_output.f90:247-264): the surface nodes
! within 20 km of the fault's extent, in node order, and their coordinates in
! surface_coor.txt<rank>.  Restated because the reference's routine re-opens
! its output unit, still connected, for every node, which this compiler's
! runtime rejects; the selection and the file are the same.
subroutine host_find_surface_nodes
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
end subroutine host_find_surface_nodes
