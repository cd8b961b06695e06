!> ISO_C_BINDING interface of the MI355X force path (C ABI: include/dlpx.h).
!>
!> A DL_POLY host Uses this module in calculate_forces (drivers.F90:573-929)
!> and calls dlpx_link_cell_pairs / dlpx_two_body_forces where it calls
!> link_cell_pairs (drivers.F90:677) and two_body_forces (drivers.F90:786).
!> Positions, charges and forces are passed in place from config%parts
!> (corePart, 64-byte stride: particle.F90:13-20) with c_loc; nothing is
!> copied on the host and no host pointer is kept after a call returns.
Module dlpx_binding
  Use, Intrinsic :: iso_c_binding
  Implicit None
  Private

  Integer(c_int32_t), Parameter, Public :: DLPX_FP64 = 0, DLPX_MIXED = 1

  !> dlpx_params (include/dlpx.h)
  Type, Bind(C), Public :: dlpx_params
    Integer(c_int32_t) :: precision = DLPX_FP64
    Integer(c_int32_t) :: natms_max = 0
    Integer(c_int32_t) :: ntypes = 0
    Integer(c_int32_t) :: max_list = 0
    Real(c_double)     :: cutoff = 0.0_c_double
    Real(c_double)     :: padding = 0.0_c_double
    Real(c_double)     :: vdw_cutoff = 0.0_c_double
    Integer(c_int32_t) :: electrostatics = 0
    Integer(c_int32_t) :: nsplines = 8
    Integer(c_int32_t) :: kvec(3) = 0
    Real(c_double)     :: alpha = 0.0_c_double
    Real(c_double)     :: epsilon = 1.0_c_double
  End Type dlpx_params

  !> dlpx_atoms: base addresses and one byte stride for every array
  Type, Bind(C), Public :: dlpx_atoms
    Integer(c_int32_t) :: natms = 0
    Integer(c_int32_t) :: stride = 0
    Type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr, chge = c_null_ptr
    Type(c_ptr) :: fx = c_null_ptr, fy = c_null_ptr, fz = c_null_ptr
    Type(c_ptr) :: ltype = c_null_ptr
  End Type dlpx_atoms

  !> dlpx_two_body_out: buffer(0:19) holds the slots of two_body.F90:708-727
  Type, Bind(C), Public :: dlpx_two_body_out
    Real(c_double) :: buffer(0:19)
    Real(c_double) :: stress(9)
    Real(c_double) :: engcpe, vircpe, engsrp, virsrp
  End Type dlpx_two_body_out

  Public :: dlpx_create, dlpx_destroy, dlpx_version, dlpx_set_vdw_tables, dlpx_set_erfc_tables, &
            dlpx_generate_vdw_tables, dlpx_generate_erfc_tables, dlpx_set_lrc, dlpx_compute_lrc, &
            dlpx_get_lrc, dlpx_link_cell_pairs, dlpx_vnl_check, dlpx_two_body_forces, dlpx_md_load, &
            dlpx_md_run, dlpx_md_store, dlpx_md_run_ensemble, dlpx_md_last_output, &
            dlpx_set_vdw_force_shift, dlpx_set_vdw_direct, dlpx_set_exclusions, dlpx_set_per_particle, &
            dlpx_add_per_particle, dlpx_set_frozen, dlpx_compute_lrc_frozen, dlpx_error_message

  Interface
    Integer(c_int) Function dlpx_create(params, ctx) Bind(C, name='dlpx_create')
      Import :: c_int, c_ptr, dlpx_params
      Type(dlpx_params), Intent(In) :: params
      Type(c_ptr), Intent(Out) :: ctx
    End Function dlpx_create

    Integer(c_int) Function dlpx_destroy(ctx) Bind(C, name='dlpx_destroy')
      Import :: c_int, c_ptr
      Type(c_ptr), Value :: ctx
    End Function dlpx_destroy

    Integer(c_int) Function dlpx_version() Bind(C, name='dlpx_version')
      Import :: c_int
    End Function dlpx_version

    Type(c_ptr) Function dlpx_last_error_c(ctx) Bind(C, name='dlpx_last_error')
      Import :: c_ptr
      Type(c_ptr), Value :: ctx
    End Function dlpx_last_error_c

    !> vdws%tab_potential / vdws%tab_force as (0:max_grid, nkeys), vdws%ltp
    Integer(c_int) Function dlpx_set_vdw_tables(ctx, nkeys, max_grid, tab_potential, tab_force, ltp) &
        Bind(C, name='dlpx_set_vdw_tables')
      Import :: c_int, c_int32_t, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: nkeys, max_grid
      Real(c_double), Intent(In) :: tab_potential(*), tab_force(*)
      Integer(c_int32_t), Intent(In) :: ltp(*)
    End Function dlpx_set_vdw_tables

    !> electro%erfc%table / electro%erfc_deriv%table
    Integer(c_int) Function dlpx_set_erfc_tables(ctx, nsamples, erfc, erfc_deriv) &
        Bind(C, name='dlpx_set_erfc_tables')
      Import :: c_int, c_int32_t, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: nsamples
      Real(c_double), Intent(In) :: erfc(*), erfc_deriv(*)
    End Function dlpx_set_erfc_tables

    !> params: vdws%param(1:7, key) for every pair key (7 per key)
    Integer(c_int) Function dlpx_generate_vdw_tables(ctx, nkeys, kind, params) &
        Bind(C, name='dlpx_generate_vdw_tables')
      Import :: c_int, c_int32_t, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: nkeys
      Integer(c_int32_t), Intent(In) :: kind(*)
      Real(c_double), Intent(In) :: params(*)
    End Function dlpx_generate_vdw_tables

    Integer(c_int) Function dlpx_generate_erfc_tables(ctx) Bind(C, name='dlpx_generate_erfc_tables')
      Import :: c_int, c_ptr
      Type(c_ptr), Value :: ctx
    End Function dlpx_generate_erfc_tables

    !> vdws%elrc, vdws%vlrc
    Integer(c_int) Function dlpx_set_lrc(ctx, elrc, vlrc) Bind(C, name='dlpx_set_lrc')
      Import :: c_int, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Real(c_double), Value :: elrc, vlrc
    End Function dlpx_set_lrc

    Integer(c_int) Function dlpx_compute_lrc(ctx, cell, num_type) Bind(C, name='dlpx_compute_lrc')
      Import :: c_int, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(In) :: cell(9), num_type(*)
    End Function dlpx_compute_lrc

    Integer(c_int) Function dlpx_get_lrc(ctx, elrc, vlrc) Bind(C, name='dlpx_get_lrc')
      Import :: c_int, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(Out) :: elrc, vlrc
    End Function dlpx_get_lrc

    !> link_cell_pairs (neighbours.F90:356)
    Integer(c_int) Function dlpx_link_cell_pairs(ctx, cell, atoms, npairs, ibig) &
        Bind(C, name='dlpx_link_cell_pairs')
      Import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double, dlpx_atoms
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(In) :: cell(9)
      Type(dlpx_atoms), Intent(In) :: atoms
      Integer(c_int64_t), Intent(Out) :: npairs
      Integer(c_int32_t), Intent(Out) :: ibig
    End Function dlpx_link_cell_pairs

    !> vnl_check (neighbours.F90:123)
    Integer(c_int) Function dlpx_vnl_check(ctx, atoms, update) Bind(C, name='dlpx_vnl_check')
      Import :: c_int, c_int32_t, c_ptr, dlpx_atoms
      Type(c_ptr), Value :: ctx
      Type(dlpx_atoms), Intent(In) :: atoms
      Integer(c_int32_t), Intent(Out) :: update
    End Function dlpx_vnl_check

    !> two_body_forces (two_body.F90:76)
    Integer(c_int) Function dlpx_two_body_forces(ctx, cell, atoms, out) Bind(C, name='dlpx_two_body_forces')
      Import :: c_int, c_ptr, c_double, dlpx_atoms, dlpx_two_body_out
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(In) :: cell(9)
      Type(dlpx_atoms), Intent(In) :: atoms
      Type(dlpx_two_body_out), Intent(Out) :: out
    End Function dlpx_two_body_forces

    !> Resident mode (nve.F90:127-215 steps on the device).  vx, vy, vz and
    !> mass are read with byte stride vstride (0 = the atoms' stride), e.g.
    !> c_loc(config%vxx(1)) ... c_loc(config%weight(1)) with vstride 8.
    Integer(c_int) Function dlpx_md_load(ctx, cell, atoms, vx, vy, vz, mass, vstride) Bind(C, name='dlpx_md_load')
      Import :: c_int, c_int32_t, c_ptr, c_double, dlpx_atoms
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(In) :: cell(9)
      Type(dlpx_atoms), Intent(In) :: atoms
      Type(c_ptr), Value :: vx, vy, vz, mass
      Integer(c_int32_t), Value :: vstride
    End Function dlpx_md_load

    !> energies: c_null_ptr, or c_loc(e(1,1)) of e(5, nsteps) (engcpe, vircpe,
    !> engsrp, virsrp, kinetic energy per step)
    Integer(c_int) Function dlpx_md_run(ctx, nsteps, tstep, energies, nrebuilds) Bind(C, name='dlpx_md_run')
      Import :: c_int, c_int32_t, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: nsteps
      Real(c_double), Value :: tstep
      Type(c_ptr), Value :: energies
      Integer(c_int32_t), Intent(Out) :: nrebuilds
    End Function dlpx_md_run

    !> positions and forces written with byte stride stride, velocities with
    !> vstride (0 = stride); any pointer may be c_null_ptr
    Integer(c_int) Function dlpx_md_store(ctx, x, y, z, vx, vy, vz, fx, fy, fz, stride, vstride) &
        Bind(C, name='dlpx_md_store')
      Import :: c_int, c_int32_t, c_ptr
      Type(c_ptr), Value :: ctx, x, y, z, vx, vy, vz, fx, fy, fz
      Integer(c_int32_t), Value :: stride, vstride
    End Function dlpx_md_store

    !> two-body output (energies, virials, stress) of the last step of the last
    !> dlpx_md_run that recorded energies
    Integer(c_int) Function dlpx_md_last_output(ctx, out) Bind(C, name='dlpx_md_last_output')
      Import :: c_int, c_ptr, dlpx_two_body_out
      Type(c_ptr), Value :: ctx
      Type(dlpx_two_body_out), Intent(Out) :: out
    End Function dlpx_md_last_output

    !> vdW force-shifting (vdws%l_force_shift): on = 0 or 1
    Integer(c_int) Function dlpx_set_vdw_force_shift(ctx, on) Bind(C, name='dlpx_set_vdw_force_shift')
      Import :: c_int, c_int32_t, c_ptr
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: on
    End Function dlpx_set_vdw_force_shift

    !> vdw_method direct (vdws%l_direct): on = 0 or 1, after dlpx_generate_vdw_tables
    Integer(c_int) Function dlpx_set_vdw_direct(ctx, on) Bind(C, name='dlpx_set_vdw_direct')
      Import :: c_int, c_int32_t, c_ptr
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: on
    End Function dlpx_set_vdw_direct

    !> excluded pairs: pass neigh%list_excl(0:mxexcl, 1:natms) as it is
    Integer(c_int) Function dlpx_set_exclusions(ctx, natms, mxexcl, list_excl) &
        Bind(C, name='dlpx_set_exclusions')
      Import :: c_int, c_int32_t, c_ptr
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: natms, mxexcl
      Integer(c_int32_t), Intent(In) :: list_excl(0:mxexcl, *)
    End Function dlpx_set_exclusions

    !> frozen atoms: pass config%lfrzn(1:natms)
    Integer(c_int) Function dlpx_set_frozen(ctx, natms, lfrzn) Bind(C, name='dlpx_set_frozen')
      Import :: c_int, c_int32_t, c_ptr
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: natms
      Integer(c_int32_t), Intent(In) :: lfrzn(*)
    End Function dlpx_set_frozen

    !> dlpx_compute_lrc with the number of frozen atoms of each type (vdw.F90:665-670)
    Integer(c_int) Function dlpx_compute_lrc_frozen(ctx, cell, num_type, num_frozen) &
        Bind(C, name='dlpx_compute_lrc_frozen')
      Import :: c_int, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(In) :: cell(9), num_type(*), num_frozen(*)
    End Function dlpx_compute_lrc_frozen

    !> stats%collect_pp: on = 0 or 1 for the following dlpx_two_body_forces calls
    Integer(c_int) Function dlpx_set_per_particle(ctx, on) Bind(C, name='dlpx_set_per_particle')
      Import :: c_int, c_int32_t, c_ptr
      Type(c_ptr), Value :: ctx
      Integer(c_int32_t), Value :: on
    End Function dlpx_set_per_particle

    !> adds the last call's values into stats%pp_energy(natms), stats%pp_stress(9, natms)
    Integer(c_int) Function dlpx_add_per_particle(ctx, pp_energy, pp_stress) Bind(C, name='dlpx_add_per_particle')
      Import :: c_int, c_ptr, c_double
      Type(c_ptr), Value :: ctx
      Real(c_double), Intent(InOut) :: pp_energy(*), pp_stress(9, *)
    End Function dlpx_add_per_particle

    !> dlpx_md_run over many independent contexts in lockstep
    Integer(c_int) Function dlpx_md_run_ensemble(ctxs, nctx, nsteps, tsteps, nrebuilds) &
        Bind(C, name='dlpx_md_run_ensemble')
      Import :: c_int, c_int32_t, c_ptr, c_double
      Type(c_ptr), Intent(In) :: ctxs(*)
      Integer(c_int32_t), Value :: nctx, nsteps
      Real(c_double), Intent(In) :: tsteps(*)
      Integer(c_int32_t), Intent(Out) :: nrebuilds(*)
    End Function dlpx_md_run_ensemble
  End Interface

Contains

  !> dlpx_last_error as a Fortran string
  Function dlpx_error_message(ctx) Result(msg)
    Type(c_ptr), Value :: ctx
    Character(Len=256) :: msg
    Character(Kind=c_char), Pointer :: s(:)
    Type(c_ptr) :: p
    Integer :: i
    Interface
      Integer(c_size_t) Function c_strlen(str) Bind(C, name='strlen')
        Import :: c_size_t, c_ptr
        Type(c_ptr), Value :: str
      End Function c_strlen
    End Interface

    msg = ' '
    p = dlpx_last_error_c(ctx)
    If (.not. c_associated(p)) Return
    Call c_f_pointer(p, s, [Int(c_strlen(p))])
    Do i = 1, Min(Size(s), Len(msg))
      msg(i:i) = s(i)
    End Do
  End Function dlpx_error_message

End Module dlpx_binding
