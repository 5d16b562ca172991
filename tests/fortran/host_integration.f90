!
! host_integration.f90 -- the unchanged MSTM Fortran modules driving the GPU
! through the ISO_C_BINDING shim (mstm_amd/fortran).  The cluster state is
! built by the reference's own setup calls (as main_calling_program does);
! then the reference CPU procedures and their hip_* replacements run on the
! same state and the results are compared.  Prints one line per check and
! exits with status 1 on a mismatch.  Usage: host_integration <posfile> <n>
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
      character(len=256) :: posfile,arg
      integer :: n,i,iref,igpu,istat_r,istat_g,mi_r,mi_g,bad
      real(8) :: line(6),err_r,err_g,d1,d2,d3
      complex(8), allocatable :: amnp_r(:,:),amnp_g(:,:),ain(:,:),ao_r(:,:),ao_g(:,:),b(:),x_r(:),x_g(:)
      real(8), allocatable :: q_r(:,:,:),q_g(:,:,:)
      call get_command_argument(1,posfile)
      call get_command_argument(2,arg)
      read(arg,*) n
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
      if(bad.ne.0) then
         write(*,'(a)') 'HOST INTEGRATION FAILED'
         call exit(1)
      endif
      write(*,'(a)') 'HOST INTEGRATION OK'
      contains
         subroutine observables_check()
         integer :: j,nt
         real(8) :: r0(3),rt,ct(361),ph(361),gg,s2(0:1),e1,e2,e3
         real(8), allocatable :: sm_r(:,:),sm_g(:,:),s00(:,:,:),s02(:,:,:),sp22(:,:,:),sm22(:,:,:)
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
         write(*,'(a,i4,3es12.4,f10.6)') 'observables T rel_err(merge,scatmat,s00) g: ',nt,e1,e2,e3,gg
         if(e1.gt.1.d-12.or.e2.gt.1.d-12.or.e3.gt.1.d-12) bad=1
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
      end program host_integration
