!------------------------------------------------------------------------------!
! MODULE : TYPHON_GPU
!
! Function
!   iso_c_binding interface of libtyphon_gpu.so (include/typhon_gpu.h) for the
!   Typhon Fortran driver: derived types laid out as the C structs and bind(C)
!   interfaces of every entry point.  A driver built with GPU support fills
!   ty_mesh_desc from umesh after init_world (SOURCE/MAIN/main.f90:95) and
!   replaces the body of calc_rhs / the RK stage loop with these calls.
!------------------------------------------------------------------------------!
module TYPHON_GPU

use iso_c_binding

implicit none

integer(c_int), parameter :: TY_OK = 0, TY_ERR_ARG = 1, TY_ERR_HIP = 2, TY_ERR_STATE = 3, &
                             TY_ERR_UNSUPPORTED = 4, TY_ERR_NOT_CONVERGED = 5
integer(c_int32_t), parameter :: TY_PREC_FP64 = 0, TY_PREC_MIXED = 1, TY_PREC_MIXED_F32SUM = 2
! element codes of ELEMVTEX.f90 taken by ty_meshbuild_faces
integer(c_int32_t), parameter :: TY_ELEM_TRI3 = 13, TY_ELEM_QUAD4 = 24, TY_ELEM_TETRA4 = 101, TY_ELEM_HEXA8 = 141

! umesh%boco(ib) with its type code and resolved uniform parameters
type, bind(C) :: ty_boco_desc
  integer(c_int32_t) :: type
  integer(c_int32_t) :: nface
  type(c_ptr)        :: iface          ! int32 (1:nface), 1-based face indices
  type(c_ptr)        :: ipart          ! periodic partner faces or c_null_ptr
  real(c_double)     :: param(8)
endtype ty_boco_desc

! st_ustmesh after init_world
type, bind(C) :: ty_mesh_desc
  integer(c_int32_t) :: dim, ncell_int, ncell, nface, nface_int
  type(c_ptr)        :: facecell       ! c_loc(umesh%facecell%fils(1,1)), (nface,2) column-major
  type(c_ptr)        :: face           ! c_loc(umesh%mesh%iface(1,1,1)), st_face records
  type(c_ptr)        :: centre         ! c_loc(umesh%mesh%centre(1,1,1)), v3d records
  type(c_ptr)        :: volume         ! c_loc(umesh%mesh%volume(1,1,1))
  integer(c_int32_t) :: nboco
  type(c_ptr)        :: boco           ! c_loc of a ty_boco_desc array
  integer(c_int32_t) :: cell_order     ! 0 RCM, 1 Hilbert, 2 caller order
endtype ty_mesh_desc

! st_espece and the fluid model
type, bind(C) :: ty_gas
  real(c_double)     :: gamma, r_const, prandtl
  integer(c_int32_t) :: typ_visc
  real(c_double)     :: visc_dyn, visc_kin, tref, tsuth
  integer(c_int32_t) :: typ_fluid
endtype ty_gas

! mnu_spat / mnu_time
type, bind(C) :: ty_scheme
  integer(c_int32_t)     :: sch_hyp, sch_dis
  character(kind=c_char) :: method, limiter
  integer(c_int32_t)     :: gradmeth, postlimiter
  real(c_double)         :: venkat_k
  integer(c_int32_t)     :: nstage
  type(c_ptr)            :: rk_coef    ! c_loc(deftime%rk%coef(1,1))
  real(c_double)         :: cfl
  integer(c_int32_t)     :: precision
endtype ty_scheme

! matrix-free implicit solve settings and statistics
type, bind(C) :: ty_krylov
  integer(c_int32_t) :: nkrylov, max_restart, max_outer, inner_precision, precond
  real(c_double)     :: inner_tol, outer_tol
endtype ty_krylov

type, bind(C) :: ty_krylov_stats
  integer(c_int32_t) :: outer_iters, inner_iters
  real(c_double)     :: rel_res, rhs_norm
endtype ty_krylov_stats

type, bind(C) :: ty_unsteady_stats
  integer(c_int32_t) :: forward_steps, adjoint_steps, max_snapshots, snapshots_allocated
  real(c_double)     :: ms
endtype ty_unsteady_stats

! steady cycle settings and statistics (integration_cyclezone for time_steady)
type, bind(C) :: ty_steady
  integer(c_int32_t) :: local_dt, maxit, check_every
  real(c_double)     :: cfl, cfl_max, maxres, residu_ref, wlog
  real(c_double)     :: refsca(2), refvec
endtype ty_steady

type, bind(C) :: ty_steady_stats
  integer(c_int32_t) :: iterations, converged
  real(c_double)     :: cur_res, residu_ref, cfl
endtype ty_steady_stats

interface

  integer(c_int) function ty_create(ctx, device) bind(C, name="ty_create")
    import :: c_int, c_ptr
    type(c_ptr)           :: ctx
    integer(c_int), value :: device
  endfunction

  integer(c_int) function ty_destroy(ctx) bind(C, name="ty_destroy")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  endfunction

  type(c_ptr) function ty_last_error(ctx) bind(C, name="ty_last_error")
    import :: c_ptr
    type(c_ptr), value :: ctx
  endfunction

  integer(c_int) function ty_mesh_upload(ctx, mesh) bind(C, name="ty_mesh_upload")
    import :: c_int, c_ptr, ty_mesh_desc
    type(c_ptr), value :: ctx
    type(ty_mesh_desc) :: mesh
  endfunction

  integer(c_int) function ty_set_scheme(ctx, gas, scheme) bind(C, name="ty_set_scheme")
    import :: c_int, c_ptr, ty_gas, ty_scheme
    type(c_ptr), value :: ctx
    type(ty_gas)       :: gas
    type(ty_scheme)    :: scheme
  endfunction

  integer(c_int) function ty_state_set(ctx, rho, rhoe, mom) bind(C, name="ty_state_set")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double)     :: rho(*), rhoe(*), mom(*)
  endfunction

  integer(c_int) function ty_state_get(ctx, rho, rhoe, mom) bind(C, name="ty_state_get")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value :: ctx
    real(c_double)     :: rho(*), rhoe(*), mom(*)
  endfunction

  integer(c_int) function ty_timestep(ctx, dtmax, dt) bind(C, name="ty_timestep")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double), value :: dtmax
    real(c_double)        :: dt
  endfunction

  integer(c_int) function ty_residual(ctx, dt, res_rho, res_rhoe, res_mom) bind(C, name="ty_residual")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
    real(c_double)        :: res_rho(*), res_rhoe(*), res_mom(*)
  endfunction

  integer(c_int) function ty_rk_step(ctx, dt) bind(C, name="ty_rk_step")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
  endfunction

  integer(c_int) function ty_advance(ctx, nsteps, dtmax, dt_last) bind(C, name="ty_advance")
    import :: c_int, c_int32_t, c_ptr, c_double
    type(c_ptr), value        :: ctx
    integer(c_int32_t), value :: nsteps
    real(c_double), value     :: dtmax
    real(c_double)            :: dt_last
  endfunction

  integer(c_int) function ty_sync(ctx) bind(C, name="ty_sync")
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  endfunction

  integer(c_int) function ty_timing(ctx, enable, ms, count) bind(C, name="ty_timing")
    import :: c_int, c_int32_t, c_ptr, c_double
    type(c_ptr), value        :: ctx
    integer(c_int32_t), value :: enable
    real(c_double)            :: ms(5)
    integer(c_int32_t)        :: count(5)
  endfunction

  integer(c_int) function ty_mesh_info(ctx, ncell, nface_dev, nslot, nbface_dev) bind(C, name="ty_mesh_info")
    import :: c_int, c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t) :: ncell, nface_dev, nslot, nbface_dev
  endfunction

  integer(c_int) function ty_jacobian_diag(ctx, dt, diag) bind(C, name="ty_jacobian_diag")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
    real(c_double)        :: diag(*)
  endfunction

  integer(c_int) function ty_jvp(ctx, dt, v, jv) bind(C, name="ty_jvp")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
    real(c_double)        :: v(*), jv(*)
  endfunction

  integer(c_int) function ty_linear_solve(ctx, dt, kry, b, x, stats) bind(C, name="ty_linear_solve")
    import :: c_int, c_ptr, c_double, ty_krylov, ty_krylov_stats
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
    type(ty_krylov)       :: kry
    real(c_double)        :: b(*), x(*)
    type(ty_krylov_stats) :: stats
  endfunction

  integer(c_int) function ty_implicit_step(ctx, dt, kry, stats) bind(C, name="ty_implicit_step")
    import :: c_int, c_ptr, c_double, ty_krylov, ty_krylov_stats
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
    type(ty_krylov)       :: kry
    type(ty_krylov_stats) :: stats
  endfunction

  ! the reference's assembled implicit step (build_implicit_bdlu + bdlu_gmres), deftime%implicite
  ! nkrylov / maxres / max_it; info = GMRES iterations - 1, or -1 when it did not converge
  integer(c_int) function ty_implicit_bdlu_step(ctx, dt, nkrylov, maxres, max_it, info) &
                 bind(C, name="ty_implicit_bdlu_step")
    import :: c_int, c_int32_t, c_ptr, c_double
    type(c_ptr), value        :: ctx
    real(c_double), value     :: dt
    integer(c_int32_t), value :: nkrylov
    real(c_double), value     :: maxres
    integer(c_int32_t), value :: max_it
    integer(c_int32_t)        :: info
  endfunction

  ! tangent-linear / adjoint of the first-order residual; pass c_null_ptr for absent arrays
  integer(c_int) function ty_residual_tangent(ctx, dt, du, ds, dvol, dq, dres) bind(C, name="ty_residual_tangent")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
    type(c_ptr), value    :: du, ds, dvol, dq
    real(c_double)        :: dres(*)
  endfunction

  integer(c_int) function ty_residual_adjoint(ctx, dt, lam, gu, gs, gvol, gq) bind(C, name="ty_residual_adjoint")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double), value :: dt
    real(c_double)        :: lam(*), gu(*)
    type(c_ptr), value    :: gs, gvol, gq
  endfunction

  integer(c_int) function ty_unsteady_adjoint(ctx, nsteps, dts, snaps, lamN, lam0, gq, stats) &
      bind(C, name="ty_unsteady_adjoint")
    import :: c_int, c_int32_t, c_ptr, c_double, ty_unsteady_stats
    type(c_ptr), value          :: ctx
    integer(c_int32_t), value   :: nsteps, snaps
    real(c_double)              :: dts(*), lamN(*), lam0(*)
    type(c_ptr), value          :: gq
    type(ty_unsteady_stats)     :: stats
  endfunction

  integer(c_int) function ty_refcons(ctx, out) bind(C, name="ty_refcons")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double)        :: out(3)
  endfunction

  integer(c_int) function ty_local_timestep(ctx, dt_cells, dt_min) bind(C, name="ty_local_timestep")
    import :: c_int, c_ptr, c_double
    type(c_ptr), value    :: ctx
    real(c_double)        :: dt_cells(*), dt_min
  endfunction

  integer(c_int) function ty_steady_cycle(ctx, steady, res_hist, stats) bind(C, name="ty_steady_cycle")
    import :: c_int, c_ptr, c_double, ty_steady, ty_steady_stats
    type(c_ptr), value    :: ctx
    type(ty_steady)       :: steady
    real(c_double)        :: res_hist(*)
    type(ty_steady_stats) :: stats
  endfunction

  ! volume probes: quantity is a C string (trim(q)//c_null_char); type 21 VOL_AVERAGE, 22 VOL_MIN, 23 VOL_MAX
  integer(c_int) function ty_probe_define(ctx, quantity, type, id) bind(C, name="ty_probe_define")
    import :: c_int, c_ptr, c_char, c_int32_t
    type(c_ptr), value          :: ctx
    character(kind=c_char)      :: quantity(*)
    integer(c_int32_t), value   :: type
    integer(c_int32_t)          :: id
  endfunction

  integer(c_int) function ty_probe_eval(ctx, t, nprobe, ids, results) bind(C, name="ty_probe_eval")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value          :: ctx
    real(c_double), value       :: t
    integer(c_int32_t), value   :: nprobe
    integer(c_int32_t)          :: ids(*)
    real(c_double)              :: results(*)
  endfunction

  ! ty_probe_eval with its partials in a fixed per-GPU workspace (no device allocation)
  integer(c_int) function ty_probe_eval_noalloc(ctx, t, nprobe, ids, results) bind(C, name="ty_probe_eval_noalloc")
    import :: c_int, c_ptr, c_double, c_int32_t
    type(c_ptr), value          :: ctx
    real(c_double), value       :: t
    integer(c_int32_t), value   :: nprobe
    integer(c_int32_t)          :: ids(*)
    real(c_double)              :: results(*)
  endfunction

  ! the device cell order: cell(i) = the caller's cell number of device cell i
  integer(c_int) function ty_cell_order(ctx, cell) bind(C, name="ty_cell_order")
    import :: c_int, c_ptr, c_int32_t
    type(c_ptr), value :: ctx
    integer(c_int32_t) :: cell(*)
  endfunction

  ! device mesh preprocessing: create_face_connect + calc_meshgeom (MESHCONNECT.f90:34-163, MESHGEOM.f90:421-461)
  integer(c_int) function ty_meshbuild_create(mb, device) bind(C, name="ty_meshbuild_create")
    import :: c_int, c_ptr
    type(c_ptr)           :: mb
    integer(c_int), value :: device
  endfunction

  integer(c_int) function ty_meshbuild_destroy(mb) bind(C, name="ty_meshbuild_destroy")
    import :: c_int, c_ptr
    type(c_ptr), value :: mb
  endfunction

  type(c_ptr) function ty_meshbuild_error(mb) bind(C, name="ty_meshbuild_error")
    import :: c_ptr
    type(c_ptr), value :: mb
  endfunction

  ! elemvtex: cellvtex%elem(1)%elemvtex(1:ncell, 1:nv); vertex: umesh%mesh%vertex(1:nvtex,1,1)
  integer(c_int) function ty_meshbuild_faces(mb, dim, elemtype, ncell, elemvtex, nvtex, vertex, nface, nface_int) &
      bind(C, name="ty_meshbuild_faces")
    import :: c_int, c_ptr, c_int32_t, c_double
    type(c_ptr), value         :: mb
    integer(c_int32_t), value  :: dim, elemtype, ncell, nvtex
    integer(c_int32_t)         :: elemvtex(*)
    real(c_double)             :: vertex(*)
    integer(c_int32_t)         :: nface, nface_int
  endfunction

  ! boundary faces in order of first appearance: bfacevtex(1:nbface, 1:4), bcentre(1:nbface) of v3d
  integer(c_int) function ty_meshbuild_boundary(mb, bfacevtex, bcentre) bind(C, name="ty_meshbuild_boundary")
    import :: c_int, c_ptr, c_int32_t, c_double
    type(c_ptr), value :: mb
    integer(c_int32_t) :: bfacevtex(*)
    real(c_double)     :: bcentre(*)
  endfunction

  ! border(j): boundary-list position of face nface_int + j; outputs as umesh%facecell%fils(1:nface,1:2),
  ! umesh%facevtex%fils(1:nface,1:4), umesh%mesh%iface(1:nface,1,1), centre and volume of the elements
  integer(c_int) function ty_meshbuild_finish(mb, border, facecell, facevtex, face, centre, volume) &
      bind(C, name="ty_meshbuild_finish")
    import :: c_int, c_ptr, c_int32_t, c_double
    type(c_ptr), value :: mb
    integer(c_int32_t) :: border(*), facecell(*), facevtex(*)
    real(c_double)     :: face(*), centre(*), volume(*)
  endfunction

endinterface

endmodule TYPHON_GPU
