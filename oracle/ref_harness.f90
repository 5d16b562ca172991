!
! ref_harness.f90 -- full-precision dump harness around the reference MSTM
! modules (december2023/: intrinsics, mpidefs-serial, modules, fft_translation,
! scatprops, solver), which oracle/Makefile compiles unmodified from where they
! lie.  TEST INFRASTRUCTURE ONLY: it writes golden vectors and times the
! reference for the CPU baseline.  It replays the setup order of
! main_calling_program (mstm-input-37.f90:1060-1262) for a cluster read from
! a position file, then runs one job:
!   mode='setup'    orders, offsets, Mie coefficients
!   mode='matvec'   sphereinteraction on two RHS read from infile
!                   (con_tran = F,T; mie_mult = T,T), plus the bare
!                   external_to_external_expansion of the same RHS
!   mode='solve'    fixedorsoln (BiCG; method='d': the LU branch) + qtotcalc
!   mode='special'  rotcoef / axialtrancoefrecurrence / gentranmatrix samples
! With observe=.true., 'solve' continues as a single-origin run of
! main_calling_program does (mstm-input-37.f90:1159-1197, 1503-1535):
! translation orders about the mean sphere position (tranordertest),
! merge_to_common_origin of both polarizations, scatteringmatrix on the
! default 1-degree grid theta = -180..180 and fosmexpansion.
!   mode='tmatrix'  the same translation orders, then tmatrix_solution
!                   (mstm-solver-8.f90:22-332, iterative, serial): the
!                   T-matrix file it writes, its per-order log and the
!                   per-sphere efficiency sums (tcap > 0 caps the order);
!                   ranorient=.true. then runs ranorientscatmatrix on that
!                   file (scattering-matrix expansion sm, smcf)
! All outputs are raw little-endian stream files <outprefix>_<name>.bin plus
! a small text summary <outprefix>_info.txt.
!
      program ref_harness
      use mpidefs
      use intrinsics
      use numconstants
      use specialfuncs
      use surface_subroutines
      use spheredata
      use mie
      use translation
      use fft_translation
      use scatprops
      use solver
      implicit none
      character(len=16) :: mode
      character(len=256) :: posfile,infile,outprefix,jobfile
      integer :: nsphere,niter,node_order_in,nbr_model,i,n,ios,istat,maxiter,ncol,reps,r,p
! idir: incidence from below (1) or above (2), the dir argument of fixedorsoln
      integer :: idir
      real(8) :: mie_eps,eps,cell_vf,alpha,sinc,vr,maxerr,t0,t1,rmed
      real(8) :: line(6),tmin(3),tmax(3),xg
      logical :: use_fft,store_tables,observe
      real(8) :: teps,conveps
      integer :: tcap
      character(len=1) :: method
      logical :: ranorient
! excite_frac > 0: excited_spheres = distance from the mean position <=
! excite_frac times the largest (an excitation_radius run,
! mstm-input-37.f90:1264-1294); dspec / dcell_in: node_selection's
! d_specified; tbox_pad > 0: target box = sphere-centre bounds widened by
! tbox_pad on every side (a random-configuration target box)
      real(8) :: excite_frac,dcell_in,tbox_pad,rex(3),rexmax
      logical :: dspec
      logical, allocatable :: exsph(:)
      character(len=30) :: tfile30
      character(len=128) :: tfile128
      real(8), allocatable :: rsm(:,:,:),rsmcf(:,:,:)
      real(8) :: r0t(3)
      real(8), allocatable :: sqeff(:,:)
      complex(8), allocatable :: ain(:,:),aout(:,:),gout(:,:),amnp(:,:),pmnp0(:,:),pmnpan(:)
      real(8), allocatable :: qeff(:,:,:),qvabs(:,:)
      real(8) :: qtot(3,3)
      namelist /job/ mode,posfile,nsphere,mie_eps,use_fft,cell_vf,node_order_in,nbr_model, &
         eps,niter,alpha,sinc,infile,outprefix,reps,store_tables,rmed,observe,teps,conveps,tcap,method,ranorient, &
         excite_frac,dspec,dcell_in,tbox_pad,idir

      mode='setup'
      posfile='cluster.pos'
      infile='ain.bin'
      outprefix='out'
      nsphere=0
      mie_eps=1.d-6
      use_fft=.false.
      cell_vf=0.d0
      node_order_in=-1
      nbr_model=2
      eps=1.d-6
      niter=10000
      alpha=0.d0
      sinc=0.d0
      reps=1
      store_tables=.false.
      rmed=1.d0
      observe=.false.
      teps=1.d-5
      conveps=1.d-6
      tcap=0
      method='i'
      ranorient=.false.
      excite_frac=0.d0
      dspec=.false.
      dcell_in=0.d0
      tbox_pad=0.d0
      idir=1
      call get_command_argument(1,jobfile)
      open(10,file=trim(jobfile),status='old')
      read(10,nml=job)
      close(10)

      call mstm_mpi(mpi_command='init')
      number_spheres=nsphere
      allocate(sphere_radius(nsphere),sphere_position(3,nsphere),sphere_ref_index(2,0:nsphere), &
         host_sphere(nsphere),number_field_expansions(nsphere))
      open(11,file=trim(posfile),status='old')
      do i=1,nsphere
         read(11,*) line(1:6)
         sphere_position(:,i)=line(1:3)
         sphere_radius(i)=line(4)
         sphere_ref_index(1,i)=dcmplx(line(5),line(6))
         sphere_ref_index(2,i)=sphere_ref_index(1,i)
      enddo
      close(11)
      layer_ref_index(0)=dcmplx(rmed,0.d0)
      sphere_ref_index(:,0)=layer_ref_index(0)
      store_translation_matrix=store_tables
      plane_surface_present=.false.
      call plane_boundary_initialization()
      call findhostspheres()
      call sphere_layer_initialization()
      call miecoefcalc(mie_eps)
      call init(max_mie_order)
      call incident_field_initialization(alpha,sinc,idir)
      fft_translation_option=use_fft
      if(use_fft) then
         cell_volume_fraction=cell_vf
         tmin=sphere_min_position-tbox_pad
         tmax=sphere_max_position+tbox_pad
         call clear_fft_matrix(clear_h=.true.)
         neighbor_node_model=nbr_model
         if(dspec) d_cell=dcell_in
         call node_selection(cell_volume_fraction,target_min=tmin,target_max=tmax, &
            d_specified=dspec,local_host=0)
         if(node_order_in.le.0) then
            node_order=-node_order_in+ceiling(d_cell)
         else
            node_order=node_order_in
         endif
      endif

      open(20,file=trim(outprefix)//'_info.txt')
      write(20,'(a,i10)') 'nsphere ',nsphere
      write(20,'(a,i10)') 'neqns ',number_eqns
      write(20,'(a,i10)') 'max_mie_order ',max_mie_order
      if(use_fft) then
         write(20,'(a,3i8)') 'cell_dim ',cell_dim
         write(20,'(a,es25.17)') 'd_cell ',d_cell
         write(20,'(a,i8)') 'node_order ',node_order
         write(20,'(a,es25.17)') 'cell_volume_fraction ',cell_volume_fraction
      endif
      call dump_i('order',sphere_order,nsphere)
      call dump_i('offset',sphere_offset,nsphere+1)
      call dump_z('anmie',an_mie,size(an_mie))

      if(trim(mode).eq.'matvec') then
         allocate(ain(number_eqns,2),aout(number_eqns,2),gout(number_eqns,2))
         open(12,file=trim(infile),access='stream',form='unformatted',status='old')
         read(12) ain
         close(12)
         t0=mstm_mpi_wtime()
         do r=1,reps
            call sphereinteraction(number_eqns,2,ain,aout,initial_run=(r.eq.1), &
               con_tran=(/.false.,.true./),mie_mult=(/.true.,.true./))
         enddo
         t1=mstm_mpi_wtime()
         write(20,'(a,es25.17)') 'matvec_seconds ',(t1-t0)/dble(reps)
         call dump_z('aout',aout,2*number_eqns)
         if(.not.use_fft) then
            call external_to_external_expansion(number_eqns,2,ain,gout, &
               con_tran=(/.false.,.true./))
            call dump_z('hout',gout,2*number_eqns)
         endif
      elseif(trim(mode).eq.'solve') then
         allocate(amnp(number_eqns,2),qeff(3,3,nsphere),qvabs(3,nsphere))
         vr=0.d0
         do i=1,nsphere
            vr=vr+sphere_radius(i)**3
         enddo
         xg=vr**.333333
         t0=mstm_mpi_wtime()
         if(use_fft) then
!
! fft_external_to_external_expansion sizes its saved node arrays by the nrhs
! of its initial_run call (fft_translation-5.f90:190-194).  cbicg makes that
! call with nrhs=1 and then applies 2-RHS products, writing past the arrays.
! Here the initial call is made with nrhs=2, then the serial iterative branch
! of fixedorsoln (mstm-solver-8.f90:468-525, 553-564) is replayed with
! cbicg(initialize_solver=.false.), so every later call fits the allocation.
!
            allocate(ain(number_eqns,2),aout(number_eqns,2),pmnp0(number_eqns,2),pmnpan(number_eqns))
            ain=(0.d0,0.d0)
            call sphereinteraction(number_eqns,2,ain,aout,initial_run=.true., &
               con_tran=(/.false.,.true./))
            call sphereplanewavecoef(alpha,sinc,idir,pmnp0)
            do p=1,2
               istat=0
               call multmiecoeffmult(number_eqns,1,1,pmnp0(:,p),pmnpan)
               amnp(:,p)=pmnpan
               call cbicg(niter,eps,pmnpan,amnp(:,p),0,maxiter,maxerr,initialize_solver=.false.)
               if(maxiter.gt.niter.or.maxerr.gt.eps) istat=1
            enddo
            call qefficiencyfactors(nsphere,2,amnp,pmnp0,qeff)
         else
            allocate(exsph(nsphere))
            exsph=.true.
            if(excite_frac.gt.0.d0) then
               rex=0.d0
               do i=1,nsphere
                  rex=rex+sphere_position(:,i)
               enddo
               rex=rex/dble(nsphere)
               rexmax=0.d0
               do i=1,nsphere
                  rexmax=max(rexmax,sqrt(sum((sphere_position(:,i)-rex)**2)))
               enddo
               do i=1,nsphere
                  exsph(i)=sqrt(sum((sphere_position(:,i)-rex)**2)).le.excite_frac*rexmax
               enddo
               write(20,'(a,i10)') 'excited ',count(exsph)
            endif
            call fixedorsoln(alpha,sinc,idir,eps,niter,amnp,qeff,3,maxerr,maxiter,0,istat,excited_spheres=exsph, &
               solution_method=method,initialize_solver=.true.)
         endif
         t1=mstm_mpi_wtime()
         call qtotcalc(nsphere,3,xg,qeff,qvabs,qtot)
         write(20,'(a,i10)') 'iterations ',maxiter
         write(20,'(a,es25.17)') 'maxerr ',maxerr
         write(20,'(a,i10)') 'istat ',istat
         write(20,'(a,es25.17)') 'solve_seconds ',t1-t0
         write(20,'(a,es25.17)') 'cross_section_radius ',xg
         write(20,'(a,9es25.17)') 'qtot ',qtot
         call dump_z('amnp',amnp,2*number_eqns)
         call dump_d('qeff',qeff,9*nsphere)
         call dump_d('qtot',qtot,9)
         if(observe) call observables()
      elseif(trim(mode).eq.'tmatrix') then
         allocate(sqeff(3,nsphere))
! vol_radius as main_calling_program sets it (mstm-input-37.f90:1205-1211)
         vr=0.d0
         do i=1,nsphere
            vr=vr+sphere_radius(i)**3
         enddo
         vol_radius=vr**.333333
         write(20,'(a,es25.17)') 'vol_radius ',vol_radius
         call set_orders(r0t)
! tcap > 0 truncates the T-matrix order (a bounded sample for timing)
         if(tcap.gt.0) t_matrix_order=min(t_matrix_order,tcap)
         write(20,'(a,i10)') 't_matrix_order_in ',t_matrix_order
         write(20,'(a,3es25.17)') 'cluster_origin ',r0t
         call dump_i('torder',translation_order,nsphere)
         run_print_unit=21
         open(21,file=trim(outprefix)//'_tmlog.txt')
! tmatrix_solution writes its T-matrix file on unit 20: the summary is
! closed around the call and reopened for appending
         close(20)
         t0=mstm_mpi_wtime()
! t_matrix_file is character*128 in tmatrix_solution: pass a blank-padded
! variable of that length
         tfile128=trim(outprefix)//'_tmat.dat'
         call tmatrix_solution(solution_method='i',solution_eps=eps,convergence_eps=conveps, &
            max_iterations=niter,t_matrix_file=tfile128, &
            sphere_qeff=sqeff,solution_status=istat)
         t1=mstm_mpi_wtime()
         close(21)
         open(20,file=trim(outprefix)//'_info.txt',access='append')
         write(20,'(a,i10)') 't_matrix_order ',t_matrix_order
         write(20,'(a,i10)') 'istat ',istat
         write(20,'(a,es25.17)') 'solve_seconds ',t1-t0
         call dump_d('sqeff',sqeff,3*nsphere)
         if(ranorient) then
! ranorientscatmatrix (mstm-scatprops-26.f90:1963-2530) on the file just
! written (run in the output directory: the file name is character*30)
            allocate(rsm(4,4,0:2*t_matrix_order),rsmcf(4,4,0:2*t_matrix_order))
            rsm=0.d0
            rsmcf=0.d0
            cross_section_radius=vol_radius
            tfile30='out_tmat.dat'
            t0=mstm_mpi_wtime()
            call ranorientscatmatrix(tfile30,rsm,rsmcf,keep_quiet=.true.)
            t1=mstm_mpi_wtime()
            write(20,'(a,es25.17)') 'ranorient_seconds ',t1-t0
            call dump_d('rsm',rsm,size(rsm))
            call dump_d('rsmcf',rsmcf,size(rsmcf))
         endif
      elseif(trim(mode).eq.'special') then
         call special_dumps()
      endif
      close(20)

      contains

         subroutine dump_z(name,z,nz)
         character(*) :: name
         integer :: nz
         complex(8) :: z(nz)
         open(30,file=trim(outprefix)//'_'//name//'.bin',access='stream',form='unformatted', &
            status='replace')
         write(30) z
         close(30)
         end subroutine dump_z

         subroutine dump_d(name,d,nd)
         character(*) :: name
         integer :: nd
         real(8) :: d(nd)
         open(30,file=trim(outprefix)//'_'//name//'.bin',access='stream',form='unformatted', &
            status='replace')
         write(30) d
         close(30)
         end subroutine dump_d

         subroutine dump_i(name,iv,ni)
         character(*) :: name
         integer :: ni,iv(ni)
         open(30,file=trim(outprefix)//'_'//name//'.bin',access='stream',form='unformatted', &
            status='replace')
         write(30) iv
         close(30)
         end subroutine dump_i
!
! post-solve observables of a single-origin run (see the header)
!
!
! cluster origin, translation orders and t_matrix_order as
! main_calling_program sets them (mstm-input-37.f90:1159-1197)
!
         subroutine set_orders(r0)
         integer :: j
         real(8) :: r0(3),rt
         complex(8) :: rimed(2)
         r0=0.d0
         do j=1,nsphere
            r0=r0+sphere_position(:,j)
         enddo
         r0=r0/dble(nsphere)
         cluster_origin=r0
         if(allocated(translation_order)) deallocate(translation_order)
         allocate(translation_order(nsphere))
         translation_order(1:nsphere)=sphere_order(1:nsphere)
         t_matrix_order=min(max_mie_order,max_t_matrix_order)
         do j=1,nsphere
            call exteriorrefindex(j,rimed)
            rt=sqrt(sum((sphere_position(:,j)-r0(:))**2))
            call tranordertest(rt,rimed(1),sphere_order(j),teps,translation_order(j))
            translation_order(j)=min(translation_order(j),max_t_matrix_order)
            t_matrix_order=max(t_matrix_order,translation_order(j))
         enddo
         end subroutine set_orders

         subroutine observables()
         integer :: j,nt,nsm
         real(8) :: r0(3),ct,ph
         real(8), allocatable :: sm(:,:),s00(:,:,:),s02(:,:,:),sp22(:,:,:),sm22(:,:,:)
         complex(8) :: sa(4)
         complex(8), allocatable :: a0(:,:)
         call set_orders(r0)
         nt=t_matrix_order
         allocate(a0(2*nt*(nt+2),2))
         a0=(0.d0,0.d0)
         do j=1,2
            call merge_to_common_origin(nt,amnp(:,j),a0(:,j),origin_position=r0,merge_procs=.true.)
         enddo
         allocate(sm(16,-180:180))
         do j=-180,180
            ct=cos(pi/180.d0*dble(j))
            if(ct.eq.1.d0) ct=0.9999999d0
            if(ct.eq.-1.d0) ct=-0.9999999d0
            if(j.lt.0) then
               ph=alpha+pi
            else
               ph=alpha
            endif
            call scatteringmatrix(a0,nt,ct,ph,sa,sm(:,j),rotate_plane=.false.,normalize_s11=.false.)
         enddo
         allocate(s00(4,4,0:2*nt),s02(4,4,0:2*nt),sp22(4,4,0:2*nt),sm22(4,4,0:2*nt))
         call fosmexpansion(nt,a0,s00,s02,sp22,sm22)
         write(20,'(a,i10)') 't_matrix_order ',nt
         write(20,'(a,3es25.17)') 'cluster_origin ',r0
         call dump_i('torder',translation_order,nsphere)
         call dump_z('amnp0',a0,size(a0))
         nsm=size(sm)
         call dump_d('scatmat',sm,nsm)
         call dump_d('s00',s00,size(s00))
         call dump_d('s02',s02,size(s02))
         call dump_d('sp22',sp22,size(sp22))
         call dump_d('sm22',sm22,size(sm22))
         end subroutine observables
!
! special functions at fixed sample points: Wigner d, axial coefficients
! (both types, unequal orders) and the dense translation matrix.
!
         subroutine special_dumps()
         integer :: nt,ns,td
         real(8) :: dc(-4:4,0:48),xp(3)
         complex(8) :: ri(2),ac(2000),gm(48,35,2)
         call rotcoef(0.3d0,4,6,dc)
         call dump_d('rotcoef',dc,size(dc))
         ri=(1.d0,0.d0)
         nt=6
         ns=4
         td=atcdim(nt,ns)
         write(20,'(a,i8)') 'atcdim64 ',td
         call axialtrancoefrecurrence(3,7.25d0,ri,nt,ns,td,ac)
         call dump_z('axial3',ac,td)
         call axialtrancoefrecurrence(1,2.75d0,ri,nt,ns,td,ac)
         call dump_z('axial1',ac,td)
         xp=(/1.5d0,-2.25d0,3.0d0/)
         call gentranmatrix(5,6,translation_vector=xp,refractive_index=ri,ac_matrix=gm, &
            vswf_type=3,mode_s=2,mode_t=2)
         call dump_z('gentran3',gm,size(gm))
         call gentranmatrix(5,6,translation_vector=xp,refractive_index=ri,ac_matrix=gm, &
            vswf_type=1,mode_s=2,mode_t=2)
         call dump_z('gentran1',gm,size(gm))
         end subroutine special_dumps

      end program ref_harness
