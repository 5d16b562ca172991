!
! mstm_hip_c.f90 -- ISO_C_BINDING interfaces of libmstm_hip.so
! (include/mstm_hip.h).  No dependency on the MSTM modules: any Fortran host
! can bind the library with this module alone.  Complex(8) arrays are passed
! where the C side takes interleaved (re, im) doubles.
!
      module mstm_hip_c
      use iso_c_binding
      implicit none
      integer(c_int), parameter :: mstm_hip_fp64=0, mstm_hip_fp32=1
      integer(c_int), parameter :: mstm_hip_solve_plain=0, mstm_hip_solve_mixed=1, &
         mstm_hip_solve_learned=2, mstm_hip_solve_fgmres=3, mstm_hip_solve_lu=4
      interface
         integer(c_int) function mstm_hip_create(ctx,device) bind(c,name='mstm_hip_create')
         import :: c_ptr, c_int
         type(c_ptr) :: ctx
         integer(c_int), value :: device
         end function mstm_hip_create
         integer(c_int) function mstm_hip_destroy(ctx) bind(c,name='mstm_hip_destroy')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         end function mstm_hip_destroy
         type(c_ptr) function mstm_hip_last_error(ctx) bind(c,name='mstm_hip_last_error')
         import :: c_ptr
         type(c_ptr), value :: ctx
         end function mstm_hip_last_error
         integer(c_int) function mstm_hip_set_spheres(ctx,n,pos,radius,ri,ri_medium,mie_eps) &
            bind(c,name='mstm_hip_set_spheres')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: n
         real(c_double) :: pos(*),radius(*),ri(*)
         real(c_double), value :: ri_medium,mie_eps
         end function mstm_hip_set_spheres
         integer(c_int) function mstm_hip_set_mie(ctx,order,an_mie) bind(c,name='mstm_hip_set_mie')
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int) :: order(*)
         complex(c_double_complex) :: an_mie(*)
         end function mstm_hip_set_mie
         integer(c_int) function mstm_hip_number_eqns(ctx) bind(c,name='mstm_hip_number_eqns')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         end function mstm_hip_number_eqns
         integer(c_int) function mstm_hip_setup_direct(ctx,precision) bind(c,name='mstm_hip_setup_direct')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: precision
         end function mstm_hip_setup_direct
         integer(c_int) function mstm_hip_node_selection(ctx,fva,target_min,target_max,d_specified, &
            d_cell,local_host,node_order,model,cell_dim,node_order_out) bind(c,name='mstm_hip_node_selection')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double) :: fva,d_cell
         type(c_ptr), value :: target_min,target_max
         integer(c_int), value :: d_specified,local_host,node_order,model
         integer(c_int) :: cell_dim(3),node_order_out
         end function mstm_hip_node_selection
         integer(c_int) function mstm_hip_fft_grid(ctx,cell_boundary,sphere_node,number_neighbor_nodes) &
            bind(c,name='mstm_hip_fft_grid')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double) :: cell_boundary(3)
         integer(c_int) :: sphere_node(*),number_neighbor_nodes
         end function mstm_hip_fft_grid
         integer(c_int) function mstm_hip_setup_fft(ctx,precision) bind(c,name='mstm_hip_setup_fft')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: precision
         end function mstm_hip_setup_fft
         integer(c_int) function mstm_hip_sphereinteraction(ctx,nrhs,ain,aout,con_tran,mie_mult, &
            use_fft,precision,rhs_list,skip_external_translation) bind(c,name='mstm_hip_sphereinteraction')
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: nrhs,use_fft,precision,skip_external_translation
         complex(c_double_complex) :: ain(*),aout(*)
         integer(c_int) :: con_tran(*),mie_mult(*)
         type(c_ptr), value :: rhs_list
         end function mstm_hip_sphereinteraction
         integer(c_int) function mstm_hip_cbicg(ctx,niter,eps,pnp,anp,use_fft,iter,errmax) &
            bind(c,name='mstm_hip_cbicg')
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: niter,use_fft
         real(c_double), value :: eps
         complex(c_double_complex) :: pnp(*),anp(*)
         integer(c_int) :: iter
         real(c_double) :: errmax
         end function mstm_hip_cbicg
         integer(c_int) function mstm_hip_fixedorsoln(ctx,alpha,sinc,dir,eps,niter,solver,use_fft, &
            amnp,qeff,maxerr,maxiter,istat,excited_spheres) bind(c,name='mstm_hip_fixedorsoln')
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: alpha,sinc,eps
         integer(c_int), value :: dir,niter,solver,use_fft
         complex(c_double_complex) :: amnp(*)
         real(c_double) :: qeff(*),maxerr
         integer(c_int) :: maxiter,istat
         type(c_ptr), value :: excited_spheres
         end function mstm_hip_fixedorsoln
         integer(c_int) function mstm_hip_set_learned(ctx,params,nparams,hidden,nlayers,lnet) &
            bind(c,name='mstm_hip_set_learned')
         import :: c_ptr, c_int, c_float
         type(c_ptr), value :: ctx
         real(c_float) :: params(*)
         integer(c_int), value :: nparams,hidden,nlayers,lnet
         end function mstm_hip_set_learned
         integer(c_int) function mstm_hip_set_translation_orders(ctx,origin,torder,t) &
            bind(c,name='mstm_hip_set_translation_orders')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double) :: origin(3)
         integer(c_int) :: torder(*)
         integer(c_int), value :: t
         end function mstm_hip_set_translation_orders
         integer(c_int) function mstm_hip_merge_to_common_origin(ctx,amnp,amnp0) &
            bind(c,name='mstm_hip_merge_to_common_origin')
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex) :: amnp(*),amnp0(*)
         end function mstm_hip_merge_to_common_origin
         integer(c_int) function mstm_hip_scattering_matrix(ctx,t,amnp0,nang,ct,phi,sm) &
            bind(c,name='mstm_hip_scattering_matrix')
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: t,nang
         complex(c_double_complex) :: amnp0(*)
         real(c_double) :: ct(*),phi(*),sm(*)
         end function mstm_hip_scattering_matrix
         integer(c_int) function mstm_hip_asymmetry(ctx,t,amnp0,s00,g) bind(c,name='mstm_hip_asymmetry')
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: t
         complex(c_double_complex) :: amnp0(*)
         real(c_double) :: s00(2),g
         end function mstm_hip_asymmetry
         integer(c_int) function mstm_hip_fosmexpansion(ctx,t,amnp0,s00,s02,sp22,sm22) &
            bind(c,name='mstm_hip_fosmexpansion')
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         integer(c_int), value :: t
         complex(c_double_complex) :: amnp0(*)
         real(c_double) :: s00(*),s02(*),sp22(*),sm22(*)
         end function mstm_hip_fosmexpansion
         integer(c_int) function mstm_hip_direct_solver(ctx,pnp,anp,initialize,niter,solution_error) &
            bind(c,name='mstm_hip_direct_solver')
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex) :: pnp(*),anp(*)
         integer(c_int), value :: initialize,niter
         real(c_double) :: solution_error
         end function mstm_hip_direct_solver
         integer(c_int) function mstm_hip_tmatrix_solution(ctx,eps,niter,conveps,solver,tmat,tmat_cap, &
            sphere_qeff,qteff,order_log,t_order,maxiter,maxerr,istat) bind(c,name='mstm_hip_tmatrix_solution')
         import :: c_ptr, c_int, c_int64_t, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), value :: eps,conveps
         integer(c_int), value :: niter,solver
         integer(c_int64_t), value :: tmat_cap
         complex(c_double_complex) :: tmat(*)
         real(c_double) :: sphere_qeff(*),qteff(*),order_log(*),maxerr
         integer(c_int) :: t_order,maxiter,istat
         end function mstm_hip_tmatrix_solution
         integer(c_int) function mstm_hip_clear(ctx,clear_tables) bind(c,name='mstm_hip_clear')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: clear_tables
         end function mstm_hip_clear
      end interface
      contains
!
! message of the last failed call as a Fortran string
!
         function mstm_hip_message(ctx) result(msg)
         type(c_ptr), intent(in) :: ctx
         character(len=256) :: msg
         character(kind=c_char), pointer :: cs(:)
         integer :: i
         msg=' '
         call c_f_pointer(mstm_hip_last_error(ctx),cs,(/256/))
         do i=1,256
            if(cs(i).eq.c_null_char) exit
            msg(i:i)=cs(i)
         enddo
         end function mstm_hip_message
      end module mstm_hip_c
