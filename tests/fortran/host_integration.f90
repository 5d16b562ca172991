!
! host_integration.f90 -- the unchanged MSTM Fortran modules driving the GPU
! through the ISO_C_BINDING shim (mstm_amd/fortran).  The cluster state is
! built by the reference's own setup calls (as main_calling_program does);
! then the reference CPU procedures and their hip_* replacements run on the
! same state and the results are compared.  Prints one line per check and
! exits with status 1 on a mismatch.
!   host_integration <posfile> <n>       direct operator: fixedorsoln (full
!        and partial excitation, order of scattering), sphereinteraction
!        (rhs_list, skip_external_translation), cbicg, direct_solver before
!        and after a geometry change, observables (fosmexpansion included), tmatrix_solution
!   host_integration <posfile> <n> fft   node-grid operator: node_selection
!        (automatic and specified d_cell), sphereinteraction,
!        fft_external_to_external_expansion, cbicg, clear_fft_matrix
!
      program host_integration
      use mpidefs
      use numconstants
      use specialfuncs
      use surface_subroutines
      use spheredata
      use mie
      use translation
      use fft_translation
      use scatprops
      use solver
      use mstm_hip_solver
      implicit none
      character(len=256) :: posfile,arg,mode
      integer :: n,i,iref,igpu,istat_r,istat_g,mi_r,mi_g,bad
      real(8) :: line(6),err_r,err_g,d1,d2,d3
      complex(8), allocatable :: amnp_r(:,:),amnp_g(:,:),ain(:,:),ao_r(:,:),ao_g(:,:),b(:),x_r(:),x_g(:)
      real(8), allocatable :: q_r(:,:,:),q_g(:,:,:)
      call get_command_argument(1,posfile)
      call get_command_argument(2,arg)
      read(arg,*) n
      mode='direct'
      if(command_argument_count().ge.3) call get_command_argument(3,mode)
      call mstm_mpi(mpi_command='init')
      number_spheres=n
      allocate(sphere_radius(n),sphere_position(3,n),sphere_ref_index(2,0:n),host_sphere(n), &
         number_field_expansions(n))
      open(11,file=trim(posfile),status='old')
      do i=1,n
         read(11,*) line(1:6)
         sphere_position(:,i)=line(1:3)
         sphere_radius(i)=line(4)
         sphere_ref_index(1,i)=dcmplx(line(5),line(6))
         sphere_ref_index(2,i)=sphere_ref_index(1,i)
      enddo
      close(11)
      sphere_ref_index(:,0)=layer_ref_index(0)
      call plane_boundary_initialization()
      call findhostspheres()
      call sphere_layer_initialization()
      call miecoefcalc(1.d-6)
      call init(max_mie_order)
      call incident_field_initialization(0.d0,0.d0,1)
      bad=0
      if(trim(mode).eq.'fft') then
         call fft_checks()
         if(bad.ne.0) then
            write(*,'(a)') 'HOST INTEGRATION FAILED'
            call exit(1)
         endif
         write(*,'(a)') 'HOST INTEGRATION OK'
         stop
      endif
      ! fixedorsoln: reference CPU vs GPU
      allocate(amnp_r(number_eqns,2),amnp_g(number_eqns,2),q_r(3,3,n),q_g(3,3,n))
      call fixedorsoln(0.d0,0.d0,1,1.d-6,10000,amnp_r,q_r,3,err_r,mi_r,0,istat_r, &
         solution_method='i',initialize_solver=.true.)
      call hip_fixedorsoln(0.d0,0.d0,1,1.d-6,10000,amnp_g,q_g,3,err_g,mi_g,0,istat_g, &
         solution_method='i',initialize_solver=.true.)
      d1=maxval(abs(amnp_r-amnp_g))/maxval(abs(amnp_r))
      d2=maxval(abs(q_r-q_g))/maxval(abs(q_r))
      write(*,'(a,2i6,2es12.4)') 'fixedorsoln iterations(ref,gpu) rel_err(amnp,qeff): ',mi_r,mi_g,d1,d2
      if(mi_r.ne.mi_g.or.d1.gt.1.d-10.or.d2.gt.1.d-10) bad=1
      ! sphereinteraction on two columns (A x, A^H y)
      allocate(ain(number_eqns,2),ao_r(number_eqns,2),ao_g(number_eqns,2))
      do i=1,number_eqns
         ain(i,1)=dcmplx(sin(0.37d0*i),cos(0.11d0*i))
         ain(i,2)=dcmplx(cos(0.23d0*i),sin(0.05d0*i))
      enddo
      call sphereinteraction(number_eqns,2,ain,ao_r,initial_run=.true.,con_tran=(/.false.,.true./))
      call hip_sphereinteraction(number_eqns,2,ain,ao_g,initial_run=.true.,con_tran=(/.false.,.true./))
      d3=maxval(abs(ao_r-ao_g))/maxval(abs(ao_r))
      write(*,'(a,es12.4)') 'sphereinteraction rel_err: ',d3
      if(d3.gt.1.d-10) bad=1
      ! cbicg from the same start
      allocate(b(number_eqns),x_r(number_eqns),x_g(number_eqns))
      call multmiecoeffmult(number_eqns,1,1,ain(:,1),b)
      x_r=b
      x_g=b
      call cbicg(200,1.d-8,b,x_r,0,iref,err_r,initialize_solver=.true.)
      call hip_cbicg(200,1.d-8,b,x_g,0,igpu,err_g,initialize_solver=.true.)
      d1=maxval(abs(x_r-x_g))/maxval(abs(x_r))
      write(*,'(a,2i6,es12.4)') 'cbicg iterations(ref,gpu) rel_err: ',iref,igpu,d1
      if(iref.ne.igpu.or.d1.gt.1.d-10) bad=1
      ! single-origin observables of the reference solution: orders about the
      ! mean position as main_calling_program sets them, then the reference
      ! merge / scatteringmatrix / fosmexpansion against the hip_* forms
      call observables_check()
      call tmatrix_check()
      call excitation_check()
      call order_of_scattering_check()
      call rhs_list_check()
      call direct_solver_check()
      if(bad.ne.0) then
         write(*,'(a)') 'HOST INTEGRATION FAILED'
         call exit(1)
      endif
      write(*,'(a)') 'HOST INTEGRATION OK'
      contains
         subroutine observables_check()
         integer :: j,nt
         real(8) :: r0(3),rt,ct(361),ph(361),gg,s2(0:1),e1,e2,e3,e4
         real(8), allocatable :: sm_r(:,:),sm_g(:,:),s00(:,:,:),s02(:,:,:),sp22(:,:,:),sm22(:,:,:), &
            g00(:,:,:),g02(:,:,:),gp22(:,:,:),gm22(:,:,:)
         complex(8) :: rimed(2),sa(4)
         complex(8), allocatable :: a0_r(:,:),a0_g(:,:)
         r0=0.d0
         do j=1,n
            r0=r0+sphere_position(:,j)
         enddo
         r0=r0/dble(n)
         if(allocated(translation_order)) deallocate(translation_order)
         allocate(translation_order(n))
         t_matrix_order=min(max_mie_order,max_t_matrix_order)
         do j=1,n
            call exteriorrefindex(j,rimed)
            rt=sqrt(sum((sphere_position(:,j)-r0(:))**2))
            call tranordertest(rt,rimed(1),sphere_order(j),1.d-5,translation_order(j))
            translation_order(j)=min(translation_order(j),max_t_matrix_order)
            t_matrix_order=max(t_matrix_order,translation_order(j))
         enddo
         nt=t_matrix_order
         allocate(a0_r(2*nt*(nt+2),2),a0_g(2*nt*(nt+2),2),sm_r(16,361),sm_g(16,361))
         do j=1,2
            call merge_to_common_origin(nt,amnp_r(:,j),a0_r(:,j),origin_position=r0,merge_procs=.true.)
         enddo
         call hip_merge_to_common_origin(nt,amnp_r,a0_g,origin_position=r0)
         e1=maxval(abs(a0_r-a0_g))/maxval(abs(a0_r))
         do j=1,361
            ct(j)=cos(pi/180.d0*dble(j-181))
            if(ct(j).eq.1.d0) ct(j)=0.9999999d0
            ph(j)=0.d0
            if(j.lt.181) ph(j)=pi
            call scatteringmatrix(a0_r,nt,ct(j),ph(j),sa,sm_r(:,j),rotate_plane=.false.,normalize_s11=.false.)
         enddo
         call hip_scatteringmatrix(a0_r,nt,361,ct,ph,sm_g)
         e2=maxval(abs(sm_r-sm_g))/maxval(abs(sm_r))
         allocate(s00(4,4,0:2*nt),s02(4,4,0:2*nt),sp22(4,4,0:2*nt),sm22(4,4,0:2*nt))
         call fosmexpansion(nt,a0_r,s00,s02,sp22,sm22)
         call hip_asymmetry(a0_r,nt,s2,gg)
         e3=max(abs(s2(0)-s00(1,1,0))/abs(s00(1,1,0)),abs(s2(1)-s00(1,1,1))/abs(s00(1,1,0)))
         allocate(g00(4,4,0:2*nt),g02(4,4,0:2*nt),gp22(4,4,0:2*nt),gm22(4,4,0:2*nt))
         call hip_fosmexpansion(nt,a0_r,g00,g02,gp22,gm22)
         e4=max(maxval(abs(g00-s00))/maxval(abs(s00)),maxval(abs(g02-s02))/maxval(abs(s02)), &
            maxval(abs(gp22-sp22))/maxval(abs(sp22)),maxval(abs(gm22-sm22))/maxval(abs(sm22)))
         write(*,'(a,i4,4es12.4,f10.6)') 'observables T rel_err(merge,scatmat,s00,fosm) g: ',nt,e1,e2,e3,e4,gg
         if(e1.gt.1.d-12.or.e2.gt.1.d-12.or.e3.gt.1.d-12.or.e4.gt.1.d-12) bad=1
         end subroutine observables_check
!
! tmatrix_solution against hip_tmatrix_solution on the same state, to
! common-origin order 3 (the orders and origin of observables_check): the
! two T-matrix files, the order reached and sphere_qeff
!
         subroutine tmatrix_check()
         integer :: j,ios,st_r,st_g,h_r(2),h_g(2),nv
         real(8) :: vr,sq_r(3,n),sq_g(3,n),e1,e2,v_r(2),v_g(2),vmax
         character(len=128) :: f_r,f_g
         character(len=64) :: l_r,l_g
         vr=0.d0
         do j=1,n
            vr=vr+sphere_radius(j)**3
         enddo
         vol_radius=vr**.333333
         cluster_origin=0.d0
         do j=1,n
            cluster_origin=cluster_origin+sphere_position(:,j)
         enddo
         cluster_origin=cluster_origin/dble(n)
         f_r='tmat_ref.dat'
         f_g='tmat_gpu.dat'
         t_matrix_order=3
         call tmatrix_solution(solution_method='i',solution_eps=1.d-6,convergence_eps=1.d-6, &
            max_iterations=10000,t_matrix_file=f_r,sphere_qeff=sq_r,solution_status=st_r)
         t_matrix_order=3
         call hip_tmatrix_solution(solution_method='i',solution_eps=1.d-6,convergence_eps=1.d-6, &
            max_iterations=10000,t_matrix_file=f_g,sphere_qeff=sq_g,solution_status=st_g)
         open(31,file=f_r,status='old')
         open(32,file=f_g,status='old')
         read(31,'(2i4)') h_r
         read(32,'(2i4)') h_g
         e1=0.d0
         vmax=0.d0
         nv=0
         do
            read(31,'(a)',iostat=ios) l_r
            if(ios.ne.0) exit
            read(32,'(a)',iostat=ios) l_g
            if(ios.ne.0) then
               nv=-1
               exit
            endif
            read(l_r(2:index(l_r,')')-1),*) v_r
            read(l_g(2:index(l_g,')')-1),*) v_g
            e1=max(e1,maxval(abs(v_r-v_g)))
            vmax=max(vmax,maxval(abs(v_r)))
            nv=nv+1
         enddo
         close(31)
         close(32)
         e1=e1/vmax
         e2=maxval(abs(sq_r-sq_g))/maxval(abs(sq_r))
         write(*,'(a,2i4,i8,2es12.4)') 'tmatrix_solution order(ref,gpu) values rel_err(tmat,sphere_qeff): ', &
            h_r(1),h_g(1),nv,e1,e2
         if(h_r(1).ne.h_g(1).or.nv.le.0.or.st_r.ne.st_g.or.e1.gt.1.d-9.or.e2.gt.1.d-10) bad=1
         end subroutine tmatrix_check
!
! fixedorsoln with excited_spheres from an excitation radius about the mean
! position, as main_calling_program sets sphere_excitation_switch
! (mstm-input-37.f90:1264-1294): radius 0.8 of the largest distance
!
         subroutine excitation_check()
         logical :: ex(n)
         real(8) :: r0(3),rt(n),e1,e2
         integer :: j
         r0=0.d0
         do j=1,n
            r0=r0+sphere_position(:,j)
         enddo
         r0=r0/dble(n)
         do j=1,n
            rt(j)=sqrt(sum((sphere_position(:,j)-r0)**2))
         enddo
         ex=rt.le.0.8d0*maxval(rt)
         call fixedorsoln(0.3d0,0.4d0,1,1.d-6,10000,amnp_r,q_r,3,err_r,mi_r,0,istat_r, &
            excited_spheres=ex,solution_method='i',initialize_solver=.true.)
         call hip_fixedorsoln(0.3d0,0.4d0,1,1.d-6,10000,amnp_g,q_g,3,err_g,mi_g,0,istat_g, &
            excited_spheres=ex,solution_method='i',initialize_solver=.false.)
         e1=maxval(abs(amnp_r-amnp_g))/maxval(abs(amnp_r))
         e2=maxval(abs(q_r-q_g))/maxval(abs(q_r))
         write(*,'(a,i4,a,i4,2i6,2es12.4)') 'fixedorsoln excited ',count(ex),' of',n, &
            mi_r,mi_g,e1,e2
         if(mi_r.ne.mi_g.or.istat_r.ne.istat_g.or..not.(e1.le.1.d-10.and.e2.le.1.d-10).or.count(ex).eq.n) bad=1
         end subroutine excitation_check
!
! max_iterations < 0: cbicg's order-of-scattering series, through
! fixedorsoln and through cbicg itself (from a start other than b)
!
         subroutine order_of_scattering_check()
         complex(8) :: p(number_eqns),xr(number_eqns),xg(number_eqns)
         real(8) :: e1,e2
         call fixedorsoln(0.d0,0.d0,1,1.d-8,-40,amnp_r,q_r,3,err_r,mi_r,0,istat_r, &
            solution_method='i',initialize_solver=.true.)
         call hip_fixedorsoln(0.d0,0.d0,1,1.d-8,-40,amnp_g,q_g,3,err_g,mi_g,0,istat_g, &
            solution_method='i',initialize_solver=.false.)
         e1=maxval(abs(amnp_r-amnp_g))/maxval(abs(amnp_r))
         e2=abs(err_r-err_g)/max(abs(err_r),1.d-300)
         write(*,'(a,4i6,2es12.4)') 'order of scattering fixedorsoln iterations(ref,gpu) istat(ref,gpu) rel_err: ', &
            mi_r,mi_g,istat_r,istat_g,e1,e2
         if(mi_r.ne.mi_g.or.istat_r.ne.istat_g.or..not.(e1.le.1.d-10.and.e2.le.1.d-8)) bad=1
         call multmiecoeffmult(number_eqns,1,1,ain(:,2),p)
         xr=ain(:,1)
         xg=ain(:,1)
         call cbicg(-7,1.d-30,p,xr,0,iref,err_r,initialize_solver=.true.)
         call hip_cbicg(-7,1.d-30,p,xg,0,igpu,err_g,initialize_solver=.false.)
         e1=maxval(abs(xr-xg))/maxval(abs(xr))
         write(*,'(a,2i6,2es12.4,es12.4)') 'order of scattering cbicg iterations(ref,gpu) errmax(ref,gpu) rel_err: ', &
            iref,igpu,err_r,err_g,e1
         if(iref.ne.igpu.or.err_r.ne.0.d0.or.err_g.ne.0.d0.or..not.(e1.le.1.d-10)) bad=1
         end subroutine order_of_scattering_check
!
! sphereinteraction's rhs_list and skip_external_translation
!
         subroutine rhs_list_check()
         real(8) :: e1,e2
         ao_r=(1.d0,1.d0)
         ao_g=(1.d0,1.d0)
         call sphereinteraction(number_eqns,2,ain,ao_r,rhs_list=(/.false.,.true./),con_tran=(/.false.,.true./))
         call hip_sphereinteraction(number_eqns,2,ain,ao_g,rhs_list=(/.false.,.true./),con_tran=(/.false.,.true./))
         e1=maxval(abs(ao_r-ao_g))/maxval(abs(ao_r))
         call sphereinteraction(number_eqns,2,ain,ao_r,skip_external_translation=.true.)
         call hip_sphereinteraction(number_eqns,2,ain,ao_g,skip_external_translation=.true.)
         e2=maxval(abs(ao_r-ao_g))
         write(*,'(a,2es12.4)') 'sphereinteraction rhs_list rel_err, skip_external_translation abs_err: ',e1,e2
         if(.not.(e1.le.1.d-10).or.e2.ne.0.d0.or.maxval(abs(ao_r)).ne.0.d0) bad=1
         end subroutine rhs_list_check
!
! direct_solver (LU of the dense interaction matrix) against hip_direct_solver,
! then after a geometry change (sphere 1 moved outward, the host's Mie data
! made again) with initialize_solver, and a reuse of the factors
!
         subroutine direct_solver_check()
         complex(8) :: p(number_eqns),xr(number_eqns),xg(number_eqns)
         real(8) :: e1,e2,e3,s_r,s_g
         call multmiecoeffmult(number_eqns,1,1,ain(:,1),p)
         call direct_solver(p,xr,initialize_solver=.true.,solution_error=s_r)
         call hip_direct_solver(p,xg,initialize_solver=.true.,solution_error=s_g)
         e1=maxval(abs(xr-xg))/maxval(abs(xr))
         sphere_position(:,1)=sphere_position(:,1)*1.25d0
         call sphere_layer_initialization()
         call miecoefcalc(1.d-6)
         call multmiecoeffmult(number_eqns,1,1,ain(:,2),p)
         call direct_solver(p,xr,initialize_solver=.true.,solution_error=s_r)
         call hip_direct_solver(p,xg,initialize_solver=.true.,solution_error=s_g)
         e2=maxval(abs(xr-xg))/maxval(abs(xr))
         call multmiecoeffmult(number_eqns,1,1,ain(:,1),p)
         call direct_solver(p,xr,solution_error=s_r)
         call hip_direct_solver(p,xg,solution_error=s_g)
         e3=maxval(abs(xr-xg))/maxval(abs(xr))
         write(*,'(a,3es12.4)') 'direct_solver rel_err (first, after geometry change, factors reused): ',e1,e2,e3
         if(.not.(e1.le.1.d-10.and.e2.le.1.d-10.and.e3.le.1.d-10)) bad=1
         end subroutine direct_solver_check
!
! the node-grid (FFT) operator through the shim: hip_node_selection in place
! of node_selection (automatic cell size; then, after hip_clear_fft_matrix, a
! specified d_cell 1.3 times larger).  The reference's serial FFT path sizes
! its saved node arrays by the first call's nrhs (fft_translation-5.f90:
! 190-194), so each grid starts with a 2-column initial call and the solves
! run cbicg with initialize_solver=.false..  The reference keeps a complex(4)
! node-kernel spectrum: products agree to 1e-5.
!
         subroutine fft_checks()
         real(8) :: tmin(3),tmax(3),e1,e2,e3,dsp
         integer :: pass,cd0(3)
         allocate(ain(number_eqns,2),ao_r(number_eqns,2),ao_g(number_eqns,2))
         allocate(b(number_eqns),x_r(number_eqns),x_g(number_eqns))
         do i=1,number_eqns
            ain(i,1)=dcmplx(sin(0.37d0*i),cos(0.11d0*i))
            ain(i,2)=dcmplx(cos(0.23d0*i),sin(0.05d0*i))
         enddo
         fft_translation_option=.true.
         neighbor_node_model=2
         do pass=1,2
            tmin=sphere_min_position
            tmax=sphere_max_position
            cell_volume_fraction=0.d0
            call hip_clear_fft_matrix(clear_h=.true.)
            if(pass.eq.1) then
               call hip_node_selection(cell_volume_fraction,target_min=tmin,target_max=tmax, &
                  d_specified=.false.,local_host=0)
               cd0=cell_dim
               dsp=1.3d0*d_cell
            else
               d_cell=dsp
               call hip_node_selection(cell_volume_fraction,target_min=tmin,target_max=tmax, &
                  d_specified=.true.,local_host=0)
            endif
            node_order=1+ceiling(d_cell)
            call sphereinteraction(number_eqns,2,ain,ao_r,initial_run=.true.,con_tran=(/.false.,.true./))
            call hip_sphereinteraction(number_eqns,2,ain,ao_g,initial_run=.true.,con_tran=(/.false.,.true./))
            e1=maxval(abs(ao_r-ao_g))/maxval(abs(ao_r))
            call fft_external_to_external_expansion(number_eqns,2,ain,ao_r,con_tran=(/.false.,.true./))
            call hip_fft_external_to_external_expansion(number_eqns,2,ain,ao_g,con_tran=(/.false.,.true./))
            e2=maxval(abs(ao_r-ao_g))/maxval(abs(ao_r))
            call multmiecoeffmult(number_eqns,1,1,ain(:,1),b)
            x_r=b
            x_g=b
            call cbicg(10000,1.d-6,b,x_r,0,iref,err_r,initialize_solver=.false.)
            call hip_cbicg(10000,1.d-6,b,x_g,0,igpu,err_g,initialize_solver=.false.)
            e3=maxval(abs(x_r-x_g))/maxval(abs(x_r))
            write(*,'(a,i2,3i4,es14.6,i4,a,3es12.4,2i6)') 'fft pass',pass,cell_dim,d_cell,node_order, &
               ' rel_err(sphereinteraction,fft_expansion,cbicg) iterations(ref,gpu): ',e1,e2,e3,iref,igpu
            if(.not.(e1.le.1.d-5.and.e2.le.1.d-5.and.e3.le.1.d-5).or.iref.ne.igpu) bad=1
         enddo
         if(all(cell_dim.eq.cd0)) bad=1  ! the specified d_cell must give another grid
         end subroutine fft_checks
      end program host_integration
