!
! mstm_hip_solver.f90 -- drop-in GPU forms of MSTM's solver entry points.
!
! Compiled against MSTM's own modules (spheredata, mie, fft_translation,
! numconstants, surface_subroutines), whose state it reads exactly as the
! Fortran code sets it.  Each procedure has the argument list of the
! reference procedure it replaces:
!   hip_cbicg               <- cbicg               (mstm-solver-8.f90:654-655)
!   hip_fixedorsoln         <- fixedorsoln         (mstm-solver-8.f90:350-352)
!   hip_sphereinteraction   <- sphereinteraction   (mstm-scatprops-26.f90:25-27)
!   hip_merge_to_common_origin <- the two merge_to_common_origin calls of
!                              main_calling_program (mstm-input-37.f90:1510-1515)
!   hip_scatteringmatrix    <- the scatteringmatrix loop of
!                              scattering_matrix_calculation (mstm-input-37.f90:3036-3062)
!   hip_asymmetry           <- s00(1,1,0:1) of fosmexpansion (mstm-scatprops-26.f90:1820)
!   hip_fosmexpansion       <- fosmexpansion       (mstm-scatprops-26.f90:1820-1951)
!   hip_tmatrix_solution    <- tmatrix_solution    (mstm-solver-8.f90:22-25)
!   hip_direct_solver       <- direct_solver       (mstm-solver-8.f90:571-572)
!   hip_node_selection      <- node_selection      (fft_translation-5.f90:654)
!   hip_clear_fft_matrix    <- clear_fft_matrix    (fft_translation-5.f90:42)
!   hip_fft_external_to_external_expansion
!                           <- fft_external_to_external_expansion
!                                                  (fft_translation-5.f90:72-74)
! The GPU context is created on first use and refilled when
! initialize_solver / initial_run is true (geometry, orders and an_mie are
! taken from the MSTM modules, so the Mie data are the host's own).
! fft_translation_option selects the node-grid operator.  Its grid is made
! from the arguments of the last hip_node_selection call (the reference's
! node_selection runs as well, so the host's module state is its own); the
! GPU grid must equal the cell_dim / d_cell of the fft_translation module,
! or the run stops with a message.
!
      module mstm_hip_solver
      use iso_c_binding
      use mstm_hip_c
      use numconstants
      use spheredata
      use mie
      use fft_translation
      use surface_subroutines
      implicit none
      type(c_ptr), save :: hip_ctx=c_null_ptr
      logical, save :: hip_loaded=.false.
      integer, save :: hip_precision=0
! arguments of the last hip_node_selection (fft_translation-5.f90:654)
      logical, save :: hip_grid_args=.false.,hip_grid_dspec=.false.,hip_grid_box=.false.
      real(8), target, save :: hip_grid_tmin(3),hip_grid_tmax(3)
      real(8), save :: hip_grid_fva=0.d0,hip_grid_d=0.d0
! firstrun defaults of fixedorsoln and direct_solver (mstm-solver-8.f90:383-387, 590-594)
      logical, save :: hip_fo_first=.true.,hip_ds_first=.true.

      contains

         subroutine hip_check(rc,where)
         integer(c_int) :: rc
         character(*) :: where
         if(rc.ne.0) then
            write(run_print_unit,'('' mstm_hip: '',a,'' failed: '',a)') where, &
               trim(mstm_hip_message(hip_ctx))
            stop
         endif
         end subroutine hip_check
!
! push the MSTM sphere state into the GPU context
!
         subroutine hip_load_state()
         integer :: i,cd(3),no
         integer(c_int) :: rc
         real(c_double), allocatable :: ri(:)
         real(c_double) :: dc
         if(.not.c_associated(hip_ctx)) then
            call hip_check(mstm_hip_create(hip_ctx,0_c_int),'create')
         endif
         if(number_host_spheres.gt.0.or.number_plane_boundaries.gt.0) then
            write(run_print_unit,'('' mstm_hip: host spheres / plane boundaries are not on the GPU path'')')
            stop
         endif
         allocate(ri(2*number_spheres))
         do i=1,number_spheres
            ri(2*i-1)=dble(sphere_ref_index(1,i))
            ri(2*i)=dimag(sphere_ref_index(1,i))
         enddo
         rc=mstm_hip_set_spheres(hip_ctx,number_spheres,sphere_position,sphere_radius,ri, &
            dble(layer_ref_index(0)),1.d-6)
         call hip_check(rc,'set_spheres')
         call hip_check(mstm_hip_set_mie(hip_ctx,sphere_order,an_mie),'set_mie')
         if(fft_translation_option) then
            call hip_grid(cd,dc,no)
            if(any(cd.ne.cell_dim).or.abs(dc-d_cell).gt.1.d-12*abs(d_cell)) then
               write(run_print_unit,'('' mstm_hip: GPU node grid'',3i5,es16.8,'' differs from the host grid'', &
                  & 3i5,es16.8,'': call hip_node_selection in place of node_selection'')') cd,dc,cell_dim,d_cell
               stop
            endif
            call hip_check(mstm_hip_setup_fft(hip_ctx,int(hip_precision,c_int)),'setup_fft')
         endif
         deallocate(ri)
         hip_loaded=.true.
         end subroutine hip_load_state
!
! the GPU node grid from the last hip_node_selection arguments (or the
! defaults of node_selection: the host's fva, the sphere-centre box, no
! specified d) and the node_order rule of main_calling_program
! (mstm-input-37.f90:1256-1260; input_node_order is the module's node_order
! as main_calling_program leaves it, so it is passed as a positive order)
!
         subroutine hip_grid(cd,dc,no)
         integer(c_int) :: cd(3),no,ds
         real(c_double) :: dc,fv
         type(c_ptr) :: pmin,pmax
         fv=cell_volume_fraction
         dc=d_cell
         ds=0
         pmin=c_null_ptr
         pmax=c_null_ptr
         if(hip_grid_args) then
            fv=hip_grid_fva
            if(hip_grid_dspec) then
               ds=1
               dc=hip_grid_d
            endif
            if(hip_grid_box) then
               pmin=c_loc(hip_grid_tmin)
               pmax=c_loc(hip_grid_tmax)
            endif
         endif
         call hip_check(mstm_hip_node_selection(hip_ctx,fv,pmin,pmax,ds,dc,0_c_int,int(node_order,c_int), &
            int(neighbor_node_model,c_int),cd,no),'node_selection')
         end subroutine hip_grid
!
! node_selection (fft_translation-5.f90:654-825) with the reference's
! argument list: runs the reference routine for the host's module state and
! keeps the arguments for the GPU grid (made at the next state load).
! local_host must be 0: spheres inside a host sphere are not on the GPU path.
!
         subroutine hip_node_selection(fva,target_min,target_max,d_specified,local_host)
         real(8) :: fva
         real(8), optional :: target_min(3),target_max(3)
         logical, optional :: d_specified
         integer, optional :: local_host
         integer :: lh
         lh=0
         if(present(local_host)) lh=local_host
         if(lh.ne.0) then
            write(run_print_unit,'('' mstm_hip: node_selection with a local host sphere is not on the GPU path'')')
            stop
         endif
         hip_grid_fva=fva
         hip_grid_dspec=.false.
         if(present(d_specified)) hip_grid_dspec=d_specified
         hip_grid_d=d_cell
         hip_grid_box=.true.
         hip_grid_tmin=sphere_min_position
         hip_grid_tmax=sphere_max_position
         if(present(target_min)) hip_grid_tmin=target_min
         if(present(target_max)) hip_grid_tmax=target_max
         hip_grid_args=.true.
         call node_selection(fva,target_min=hip_grid_tmin,target_max=hip_grid_tmax, &
            d_specified=hip_grid_dspec,local_host=lh)
         hip_loaded=.false.
         end subroutine hip_node_selection
!
! clear_fft_matrix (fft_translation-5.f90:42-70): the reference routine for
! the host's state; the GPU state is reloaded at the next call, and with
! clear_h its operator tables and node-kernel spectrum are released now
!
         subroutine hip_clear_fft_matrix(clear_h)
         logical, optional :: clear_h
         logical :: ch
         ch=.false.
         if(present(clear_h)) ch=clear_h
         call clear_fft_matrix(clear_h=ch)
         if(ch.and.c_associated(hip_ctx)) call hip_check(mstm_hip_clear(hip_ctx,1_c_int),'clear')
         hip_loaded=.false.
         end subroutine hip_clear_fft_matrix
!
! cbicg (mstm-solver-8.f90:654-949) on the GPU, FP64, iteration-identical
!
         subroutine hip_cbicg(niter,eps,pnp,anp,iterwrite,iter,errmax,initialize_solver,mpi_comm)
         integer :: niter,iterwrite,iter
         integer, optional :: mpi_comm
         logical, optional :: initialize_solver
         real(8) :: eps,errmax
         complex(8) :: pnp(number_eqns),anp(number_eqns)
         logical :: init
         integer(c_int) :: it,rc,ufft
         init=.true.
         if(present(initialize_solver)) init=initialize_solver
         if(init.or.(.not.hip_loaded)) call hip_load_state()
         ufft=0
         if(fft_translation_option) ufft=1
         rc=mstm_hip_cbicg(hip_ctx,int(niter,c_int),eps,pnp,anp,ufft,it,errmax)
         call hip_check(rc,'cbicg')
         iter=it
         end subroutine hip_cbicg
!
! network of the learned solver from a raw weight file (int32 hidden,
! nlayers, lnet, nparams, then nparams float32), as mstm_amd.learned writes it
!
         subroutine hip_load_network(fname)
         character(*) :: fname
         integer :: u,hd(4)
         real(c_float), allocatable :: p(:)
         if(.not.c_associated(hip_ctx)) then
            call hip_check(mstm_hip_create(hip_ctx,0_c_int),'create')
         endif
         open(newunit=u,file=fname,access='stream',form='unformatted',status='old')
         read(u) hd
         allocate(p(hd(4)))
         read(u) p
         close(u)
         call hip_check(mstm_hip_set_learned(hip_ctx,p,int(hd(4),c_int),int(hd(1),c_int),int(hd(2),c_int), &
            int(hd(3),c_int)),'set_learned')
         deallocate(p)
         end subroutine hip_load_network
!
! fixedorsoln (mstm-solver-8.f90:350-569); solution_method 'i' (or absent)
! is the iterative FP64 solver with cbicg's iterates, 'm' the mixed FP32/FP64
! solver, 'l' the mixed solver from the learned start (hip_load_network
! first), 'g' flexible GMRES; any other letter is the direct (LU) solution,
! as in the reference (dirsoln = solution_method .ne. 'i', :378-380)
!
         subroutine hip_fixedorsoln(alpha,sinc,dir,eps,niter,amnp,qeff,qeffdim,maxerr,maxiter, &
            iterwrite,istat,mpi_comm,excited_spheres,solution_method,initialize_solver)
         integer :: niter,istat,maxiter,iterwrite,dir,qeffdim
         integer, optional :: mpi_comm
         logical, optional :: excited_spheres(number_spheres),initialize_solver
         character*1, optional :: solution_method
         real(8) :: alpha,sinc,eps,qeff(3,qeffdim,number_spheres),maxerr
         complex(8) :: amnp(number_eqns,2)
         real(8), allocatable :: q(:,:,:)
         logical :: init
         integer(c_int) :: solver,mi,ist,ufft
         integer(c_int), allocatable, target :: ex(:)
         type(c_ptr) :: pex
         init=hip_fo_first
         if(present(initialize_solver)) init=initialize_solver
         hip_fo_first=.false.
         if(init.or.(.not.hip_loaded)) call hip_load_state()
         pex=c_null_ptr
         if(present(excited_spheres)) then
            allocate(ex(number_spheres))
            ex=0
            where(excited_spheres) ex=1
            pex=c_loc(ex)
         endif
         solver=mstm_hip_solve_plain
         if(present(solution_method)) then
            if(solution_method.eq.'m') then
               solver=mstm_hip_solve_mixed
            elseif(solution_method.eq.'l') then
               solver=mstm_hip_solve_learned
            elseif(solution_method.eq.'g') then
               solver=mstm_hip_solve_fgmres
            elseif(solution_method.ne.'i') then
               solver=mstm_hip_solve_lu
            endif
         endif
         ufft=0
         if(fft_translation_option) ufft=1
         allocate(q(3,3,number_spheres))
         call hip_check(mstm_hip_fixedorsoln(hip_ctx,alpha,sinc,int(dir,c_int),eps,int(niter,c_int), &
            solver,ufft,amnp,q,maxerr,mi,ist,pex),'fixedorsoln')
         if(allocated(ex)) deallocate(ex)
         maxiter=mi
         istat=ist
         if(qeffdim.eq.1) then
            qeff(:,1,:)=q(:,1,:)
         else
            qeff(:,1:3,:)=q(:,1:3,:)
         endif
         deallocate(q)
         end subroutine hip_fixedorsoln
!
! sphereinteraction (mstm-scatprops-26.f90:25-180) for external spheres
!
         subroutine hip_sphereinteraction(neqns,nrhs,ain,aout,initial_run,rhs_list,mpi_comm, &
            con_tran,mie_mult,fft_option,store_matrix_option,skip_external_translation)
         integer :: neqns,nrhs
         integer, optional :: mpi_comm
         logical, optional :: initial_run,rhs_list(nrhs),con_tran(nrhs),mie_mult(nrhs), &
            fft_option,store_matrix_option,skip_external_translation
         complex(8) :: ain(neqns,nrhs),aout(neqns,nrhs)
         integer(c_int) :: ct(nrhs),mm(nrhs),ufft,skip
         integer(c_int), target :: rl(nrhs)
         type(c_ptr) :: prl
         integer :: r
         if(present(initial_run)) then
            if(initial_run) call hip_load_state()
         endif
         if(.not.hip_loaded) call hip_load_state()
         ct=0
         mm=1
         if(present(con_tran)) then
            do r=1,nrhs
               if(con_tran(r)) ct(r)=1
            enddo
         endif
         if(present(mie_mult)) then
            do r=1,nrhs
               if(.not.mie_mult(r)) mm(r)=0
            enddo
         endif
         ufft=0
         if(fft_translation_option) ufft=1
         if(present(fft_option)) then
            ufft=0
            if(fft_option) ufft=1
         endif
         prl=c_null_ptr
         if(present(rhs_list)) then
            rl=0
            where(rhs_list) rl=1
            prl=c_loc(rl)
         endif
         skip=0
         if(present(skip_external_translation)) then
            if(skip_external_translation) skip=1
         endif
         call hip_check(mstm_hip_sphereinteraction(hip_ctx,int(nrhs,c_int),ain,aout,ct,mm,ufft, &
            int(hip_precision,c_int),prl,skip),'sphereinteraction')
         end subroutine hip_sphereinteraction
!
! fft_external_to_external_expansion (fft_translation-5.f90:72-336): gout = H ain
! on the node-grid operator (no Mie multiplication), H^T ain for con_tran
! columns.  The C entry returns conj(H^T conj(x)) for a con_tran column
! without Mie multiplication, so those columns go in and come out conjugated.
!
         subroutine hip_fft_external_to_external_expansion(neqns,nrhs,ain,gout, &
            store_matrix_option,initial_run,rhs_list,mpi_comm,con_tran)
         integer :: neqns,nrhs
         integer, optional :: mpi_comm
         logical, optional :: store_matrix_option,initial_run,rhs_list(nrhs),con_tran(nrhs)
         complex(8) :: ain(neqns,nrhs),gout(neqns,nrhs)
         complex(8), allocatable :: at(:,:)
         integer(c_int) :: ct(nrhs),mm(nrhs)
         integer(c_int), target :: rl(nrhs)
         type(c_ptr) :: prl
         integer :: r
         if(.not.fft_translation_option) then
            write(run_print_unit,'('' mstm_hip: fft_external_to_external_expansion needs fft_translation_option'')')
            stop
         endif
         if(present(initial_run)) then
            if(initial_run) call hip_load_state()
         endif
         if(.not.hip_loaded) call hip_load_state()
         ct=0
         if(present(con_tran)) then
            do r=1,nrhs
               if(con_tran(r)) ct(r)=1
            enddo
         endif
         mm=0
         prl=c_null_ptr
         if(present(rhs_list)) then
            rl=0
            where(rhs_list) rl=1
            prl=c_loc(rl)
         endif
         allocate(at(neqns,nrhs))
         do r=1,nrhs
            if(ct(r).ne.0) then
               at(:,r)=conjg(ain(:,r))
            else
               at(:,r)=ain(:,r)
            endif
         enddo
         call hip_check(mstm_hip_sphereinteraction(hip_ctx,int(nrhs,c_int),at,gout,ct,mm,1_c_int, &
            int(hip_precision,c_int),prl,0_c_int),'fft_external_to_external_expansion')
         do r=1,nrhs
            if(ct(r).ne.0) gout(:,r)=conjg(gout(:,r))
         enddo
         deallocate(at)
         end subroutine hip_fft_external_to_external_expansion
!
! merge_to_common_origin (mstm-scatprops-26.f90:447-559) of both
! polarizations in one call, with the host's translation_order (spheredata)
! and the common-origin order nodrt (t_matrix_order)
!
         subroutine hip_merge_to_common_origin(nodrt,amnp,amnp0,origin_position)
         integer :: nodrt
         real(8), optional :: origin_position(3)
         complex(8) :: amnp(number_eqns,2),amnp0(2*nodrt*(nodrt+2),2)
         real(c_double) :: r0(3)
         integer(c_int), allocatable :: tord(:)
         if(.not.hip_loaded) call hip_load_state()
         r0=0.d0
         if(present(origin_position)) r0=origin_position
         allocate(tord(number_spheres))
         tord=translation_order(1:number_spheres)
         call hip_check(mstm_hip_set_translation_orders(hip_ctx,r0,tord,int(nodrt,c_int)), &
            'set_translation_orders')
         call hip_check(mstm_hip_merge_to_common_origin(hip_ctx,amnp,amnp0),'merge_to_common_origin')
         deallocate(tord)
         end subroutine hip_merge_to_common_origin
!
! scatteringmatrix (mstm-scatprops-26.f90:1498-1573; rotate_plane = .false.,
! normalize_s11 = .false.) at nang directions (ct, phi): sm(16,nang)
!
         subroutine hip_scatteringmatrix(amn0,nodrt,nang,ct,phi,sm)
         integer :: nodrt,nang
         real(8) :: ct(nang),phi(nang),sm(16,nang)
         complex(8) :: amn0(0:nodrt+1,nodrt,2,2)
         if(.not.hip_loaded) call hip_load_state()
         call hip_check(mstm_hip_scattering_matrix(hip_ctx,int(nodrt,c_int),amn0,int(nang,c_int),ct,phi,sm), &
            'scattering_matrix')
         end subroutine hip_scatteringmatrix
!
! s00(1,1,0:1) of fosmexpansion (mstm-scatprops-26.f90:1820-1951) and the
! asymmetry parameter g = s00(1,1,1)/(3 s00(1,1,0))
!
         subroutine hip_asymmetry(amn0,nodrt,s00,g)
         integer :: nodrt
         real(8) :: s00(0:1),g
         complex(8) :: amn0(0:nodrt+1,nodrt,2,2)
         if(.not.hip_loaded) call hip_load_state()
         call hip_check(mstm_hip_asymmetry(hip_ctx,int(nodrt,c_int),amn0,s00,g),'asymmetry')
         end subroutine hip_asymmetry
!
! fosmexpansion (mstm-scatprops-26.f90:1820-1951) with its argument list:
! s00, s02, sp22, sm22 of the common-origin expansion amnp, sign patch and
! scale included; mpi_comm is accepted and not used (one GPU)
!
         subroutine hip_fosmexpansion(ntot,amnp,s00,s02,sp22,sm22,mpi_comm)
         integer :: ntot
         integer, optional :: mpi_comm
         real(8) :: s00(4,4,0:ntot*2),s02(4,4,0:ntot*2),sp22(4,4,0:ntot*2),sm22(4,4,0:ntot*2)
         complex(8) :: amnp(0:ntot+1,ntot,2,2)
         if(.not.hip_loaded) call hip_load_state()
         call hip_check(mstm_hip_fosmexpansion(hip_ctx,int(ntot,c_int),amnp,s00,s02,sp22,sm22),'fosmexpansion')
         end subroutine hip_fosmexpansion

!
! tmatrix_solution (mstm-solver-8.f90:22-332), iterative branch, on the GPU
! about the host's cluster_origin with its translation_order and
! t_matrix_order.  Writes the T-matrix file in the reference's format
! (:125-130, 243-256, 310-317), sets t_matrix_order to the order reached and
! returns sphere_qeff and solution_status as the reference does.
! solution_method 'm' selects the mixed FP32/FP64 solver; 'i' or absent the
! FP64 one with cbicg's iterates.  procs_per_soln and mpi_comm are accepted
! and not used (one GPU).
!
         subroutine hip_tmatrix_solution(solution_method,solution_eps,convergence_eps, &
            max_iterations,t_matrix_file,procs_per_soln,mpi_comm, &
            sphere_qeff,solution_status,sphere_excitation_list)
         character*1, optional :: solution_method
         character*128, optional :: t_matrix_file
         integer, optional :: max_iterations,procs_per_soln,mpi_comm,solution_status
         real(8), optional :: solution_eps,convergence_eps,sphere_qeff(3,number_spheres)
         logical, optional :: sphere_excitation_list(number_spheres)
         integer :: l,i,k,n,nt0
         integer(c_int) :: solver,niter,tord,mi,ist
         integer(c_int), allocatable :: tr(:)
         integer(c_int64_t) :: ntm
         real(c_double) :: eps,ceps,me,qt(3)
         real(c_double), allocatable :: sq(:,:),qlog(:,:)
         complex(c_double_complex), allocatable :: tm(:)
         character*128 :: tfile
         if(.not.hip_loaded) call hip_load_state()
         solver=mstm_hip_solve_plain
         if(present(solution_method)) then
            if(solution_method.eq.'m') then
               solver=mstm_hip_solve_mixed
            elseif(solution_method.ne.'i') then
               write(run_print_unit,'('' mstm_hip: only the iterative T-matrix solution is on the GPU path'')')
               stop
            endif
         endif
         if(present(sphere_excitation_list)) then
            if(.not.all(sphere_excitation_list)) then
               write(run_print_unit,'('' mstm_hip: sphere excitation lists are not on the GPU path'')')
               stop
            endif
         endif
         eps=1.d-6
         if(present(solution_eps)) eps=solution_eps
         ceps=1.d-6
         if(present(convergence_eps)) ceps=convergence_eps
         niter=100
         if(present(max_iterations)) niter=max_iterations
         tfile='tmatrix_temp.dat'
         if(present(t_matrix_file)) tfile=t_matrix_file
         nt0=t_matrix_order
         allocate(tr(number_spheres))
         tr=translation_order(1:number_spheres)
         call hip_check(mstm_hip_set_translation_orders(hip_ctx,cluster_origin,tr,int(nt0,c_int)), &
            'set_translation_orders')
         ntm=0
         do l=1,nt0
            ntm=ntm+4*(2*l+1)*l*(l+2)
         enddo
         allocate(tm(ntm),sq(3,number_spheres),qlog(4,nt0))
         call hip_check(mstm_hip_tmatrix_solution(hip_ctx,eps,niter,ceps,solver,tm,ntm,sq,qt,qlog,tord,mi,me,ist), &
            'tmatrix_solution')
         write(run_print_unit,'(''  n   # its  qext         qabs           error'')')
         do l=1,tord
            write(run_print_unit,'(i4,i5,3e13.5)') l,nint(qlog(1,l)),qlog(2:4,l)
         enddo
         open(20,file=tfile)
         write(20,'(2i4)') tord,tord
         i=0
         do l=1,tord
            do k=1,2*(2*l+1)
               do n=1,2*l*(l+2)
                  i=i+1
                  write(20,'(''('',e18.10,'','',e18.10,'')'')') tm(i)
               enddo
            enddo
         enddo
         close(20)
         t_matrix_order=tord
         if(present(sphere_qeff)) sphere_qeff=sq
         if(present(solution_status)) solution_status=ist
         deallocate(tr,tm,sq,qlog)
         end subroutine hip_tmatrix_solution

!
! direct_solver (mstm-solver-8.f90:571-642): LU factors of the dense
! A = I - T H on the GPU (made when initialize_solver is true or none are
! held), back substitution, number_iterations refinement steps (default 3)
!
         subroutine hip_direct_solver(pnp,anp,initialize_solver,solution_error, &
            number_iterations,solution_eps,mpi_comm)
         logical, optional :: initialize_solver
         integer, optional :: number_iterations,mpi_comm
         real(8), optional :: solution_error,solution_eps
         complex(8) :: pnp(number_eqns),anp(number_eqns)
         integer(c_int) :: init,niter
         real(c_double) :: serr
         logical :: linit
         linit=hip_ds_first
         if(present(initialize_solver)) linit=initialize_solver
         if(linit.or.(.not.hip_loaded)) call hip_load_state()
         init=0
         if(linit) then
            init=1
            hip_ds_first=.false.
         endif
         niter=3
         if(present(number_iterations)) niter=number_iterations
         call hip_check(mstm_hip_direct_solver(hip_ctx,pnp,anp,init,niter,serr),'direct_solver')
         if(present(solution_error)) solution_error=serr
         end subroutine hip_direct_solver

      end module mstm_hip_solver
