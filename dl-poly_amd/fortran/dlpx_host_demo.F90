!> A Fortran host driving the force path the way calculate_forces
!> (drivers.F90:573-929) would: config%parts-like corePart array, one
!> link_cell_pairs-equivalent call, one two_body_forces-equivalent call, with
!> forces added into parts(:)%fxx/fyy/fzz in place.
!>
!> System: 1,000-ion NaCl (5^3 rocksalt cells, lattice 5.64*1.08 A) with a
!> deterministic integer-arithmetic displacement pattern, Tosi-Fumi BHM + SPME
!> (rc 12 A, padding 0.5 A, alpha 0.28, K = 32^3, order 8).
!>
!> Then 10 resident NVE steps (dt 2 fs) from those positions, with velocities
!> vx_n,d = 0.02*(((53 n + 7 d) mod 97)/96 - 1/2) A/ps and the Na/Cl masses
!> held in separate arrays (DL_POLY's config%vxx/.../config%weight layout).
!>
!> Usage: dlpx_host_demo <precision 0|1> <output file>
!> Output (stream, little-endian doubles): engcpe, vircpe, engsrp, virsrp,
!> stress(9), buffer(0:19), then fxx, fyy, fzz of every atom; then after the
!> resident steps xxx, yyy, zzz | vx, vy, vz | fxx, fyy, fzz of every atom.
Program dlpx_host_demo
  Use, Intrinsic :: iso_c_binding
  Use dlpx_binding
  Implicit None

  !> particle.F90:13-20 layout (7 reals + 2 ints = 64 bytes)
  Type, Bind(C) :: corePart
    Real(c_double)     :: xxx, yyy, zzz
    Real(c_double)     :: fxx, fyy, fzz
    Real(c_double)     :: chge
    Integer(c_int32_t) :: pad1, pad2
  End Type corePart

  Integer, Parameter :: ncells = 5, natms = 8 * ncells**3
  Type(corePart), Allocatable, Target :: parts(:)
  Integer(c_int32_t), Allocatable, Target :: ltype(:)
  Real(c_double), Allocatable, Target :: vx(:), vy(:), vz(:), weight(:)
  Integer(c_int32_t) :: nreb
  Real(c_double), Target :: cell(9)
  Real(c_double) :: a, L, basis(3, 4), params(7, 3), num_type(2)
  Integer(c_int32_t) :: kinds(3), ibig
  Integer(c_int64_t) :: npairs
  Type(dlpx_params) :: prm
  Type(dlpx_atoms) :: atoms
  Type(dlpx_two_body_out) :: out
  Type(c_ptr) :: ctx
  Integer :: i, ix, iy, iz, b, d, n, ierr, u
  Character(Len=256) :: arg, outfile

  Call get_command_argument(1, arg)
  Read (arg, *) prm%precision
  Call get_command_argument(2, outfile)

  a = 5.64_c_double * 1.08_c_double
  L = Real(ncells, c_double) * a
  cell = 0.0_c_double
  cell(1) = L; cell(5) = L; cell(9) = L
  basis = Reshape([0.0_c_double, 0.0_c_double, 0.0_c_double, 0.0_c_double, 0.5_c_double, 0.5_c_double, &
                   0.5_c_double, 0.0_c_double, 0.5_c_double, 0.5_c_double, 0.5_c_double, 0.0_c_double], [3, 4])

  Allocate (parts(natms), ltype(natms))
  n = 0
  Do i = 1, 2                      ! Na ions first, then Cl (FIELD molecule order)
    Do ix = 0, ncells - 1
      Do iy = 0, ncells - 1
        Do iz = 0, ncells - 1
          Do b = 1, 4
            n = n + 1
            parts(n)%xxx = a * (Real(ix, c_double) + basis(1, b)) - 0.5_c_double * L
            parts(n)%yyy = a * (Real(iy, c_double) + basis(2, b)) - 0.5_c_double * L
            parts(n)%zzz = a * (Real(iz, c_double) + basis(3, b)) - 0.5_c_double * L
            If (i == 2) parts(n)%xxx = parts(n)%xxx + 0.5_c_double * a
            If (parts(n)%xxx >= 0.5_c_double * L) parts(n)%xxx = parts(n)%xxx - L
            parts(n)%chge = Merge(1.0_c_double, -1.0_c_double, i == 1)
            ltype(n) = i
          End Do
        End Do
      End Do
    End Do
  End Do
  Do n = 1, natms                  ! displacement 0.1*(((37 n + 11 d) mod 101)/100 - 1/2)
    Do d = 1, 3
      Select Case (d)
      Case (1)
        parts(n)%xxx = parts(n)%xxx + 0.1_c_double * (Real(Mod(37 * n + 11 * d, 101), c_double) / 100.0_c_double - 0.5_c_double)
      Case (2)
        parts(n)%yyy = parts(n)%yyy + 0.1_c_double * (Real(Mod(37 * n + 11 * d, 101), c_double) / 100.0_c_double - 0.5_c_double)
      Case (3)
        parts(n)%zzz = parts(n)%zzz + 0.1_c_double * (Real(Mod(37 * n + 11 * d, 101), c_double) / 100.0_c_double - 0.5_c_double)
      End Select
    End Do
    parts(n)%fxx = 0.0_c_double; parts(n)%fyy = 0.0_c_double; parts(n)%fzz = 0.0_c_double
  End Do

  prm%natms_max = natms
  prm%ntypes = 2
  prm%cutoff = 12.0_c_double
  prm%padding = 0.5_c_double
  prm%vdw_cutoff = 12.0_c_double
  prm%electrostatics = 1
  prm%nsplines = 8
  prm%kvec = [32, 32, 32]
  prm%alpha = 0.28_c_double
  prm%epsilon = 1.0_c_double
  ierr = dlpx_create(prm, ctx)
  If (ierr /= 0) Stop 'dlpx_create failed'

  ! Tosi-Fumi BHM, keys (Na,Na) (Na,Cl) (Cl,Cl), internal units
  kinds = [5, 5, 5]
  ! vdws%param(1:7, key), the reference's own layout (max_param = 7)
  params = 0.0_c_double
  params(1:5, 1) = [2544.4_c_double, 3.15457_c_double, 2.340_c_double, 10117.0_c_double, 4818.0_c_double]
  params(1:5, 2) = [2035.5_c_double, 3.15457_c_double, 2.755_c_double, 67448.0_c_double, 83708.0_c_double]
  params(1:5, 3) = [1526.6_c_double, 3.15457_c_double, 3.170_c_double, 698570.0_c_double, 1403160.0_c_double]
  Call check(dlpx_generate_vdw_tables(ctx, 3_c_int32_t, kinds, params))
  Call check(dlpx_generate_erfc_tables(ctx))
  num_type = [Real(natms / 2, c_double), Real(natms / 2, c_double)]
  Call check(dlpx_compute_lrc(ctx, cell, num_type))

  atoms%natms = natms
  atoms%stride = Int(c_sizeof(parts(1)), c_int32_t)
  atoms%x = c_loc(parts(1)%xxx); atoms%y = c_loc(parts(1)%yyy); atoms%z = c_loc(parts(1)%zzz)
  atoms%chge = c_loc(parts(1)%chge)
  atoms%fx = c_loc(parts(1)%fxx); atoms%fy = c_loc(parts(1)%fyy); atoms%fz = c_loc(parts(1)%fzz)
  atoms%ltype = c_loc(ltype(1))

  Call check(dlpx_link_cell_pairs(ctx, cell, atoms, npairs, ibig))
  Call check(dlpx_two_body_forces(ctx, cell, atoms, out))
  Write (*, '(a,i0,a,i0)') 'pairs ', npairs, ' ibig ', ibig
  Write (*, '(a,4es24.16)') 'engcpe vircpe engsrp virsrp ', out%engcpe, out%vircpe, out%engsrp, out%virsrp

  Open (Newunit=u, File=Trim(outfile), Access='stream', Form='unformatted', Status='replace')
  Write (u) out%engcpe, out%vircpe, out%engsrp, out%virsrp, out%stress, out%buffer
  Write (u) (parts(n)%fxx, parts(n)%fyy, parts(n)%fzz, n=1, natms)

  ! resident steps: velocities and masses in their own arrays (stride 8)
  Allocate (vx(natms), vy(natms), vz(natms), weight(natms))
  Do n = 1, natms
    vx(n) = 0.02_c_double * (Real(Mod(53 * n + 7, 97), c_double) / 96.0_c_double - 0.5_c_double)
    vy(n) = 0.02_c_double * (Real(Mod(53 * n + 14, 97), c_double) / 96.0_c_double - 0.5_c_double)
    vz(n) = 0.02_c_double * (Real(Mod(53 * n + 21, 97), c_double) / 96.0_c_double - 0.5_c_double)
    weight(n) = Merge(22.98977_c_double, 35.453_c_double, ltype(n) == 1)
  End Do
  Call check(dlpx_md_load(ctx, cell, atoms, c_loc(vx(1)), c_loc(vy(1)), c_loc(vz(1)), c_loc(weight(1)), &
                          Int(c_sizeof(vx(1)), c_int32_t)))
  Call check(dlpx_md_run(ctx, 10_c_int32_t, 0.002_c_double, c_null_ptr, nreb))
  Call check(dlpx_md_store(ctx, c_loc(parts(1)%xxx), c_loc(parts(1)%yyy), c_loc(parts(1)%zzz), &
                           c_loc(vx(1)), c_loc(vy(1)), c_loc(vz(1)), &
                           c_loc(parts(1)%fxx), c_loc(parts(1)%fyy), c_loc(parts(1)%fzz), &
                           atoms%stride, Int(c_sizeof(vx(1)), c_int32_t)))
  Write (*, '(a,i0)') 'resident steps done, list rebuilds ', nreb
  Write (u) (parts(n)%xxx, parts(n)%yyy, parts(n)%zzz, n=1, natms)
  Write (u) (vx(n), vy(n), vz(n), n=1, natms)
  Write (u) (parts(n)%fxx, parts(n)%fyy, parts(n)%fzz, n=1, natms)
  Close (u)
  ierr = dlpx_destroy(ctx)

Contains

  Subroutine check(code)
    Integer(c_int), Intent(In) :: code
    If (code /= 0) Then
      Write (*, '(a,i0,2a)') 'dlpx error ', code, ': ', Trim(dlpx_error_message(ctx))
      Stop 1
    End If
  End Subroutine check

End Program dlpx_host_demo
