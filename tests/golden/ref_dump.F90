! Test infrastructure only (fixture generation in the build container; never
! shipped, never run on a GPU).  Linked with the compiled DL_POLY reference
! objects (oracle/_ref/obj) by tests/golden/build_ref_dump.py: after the
! reference's first two_body_forces call (two_body.F90:76, called from
! calculate_forces, drivers.F90:786) it writes, in binary and at full double
! precision, what REVCON and REVIVE carry only partly:
!   natms; per atom ltg(i) (int32) and the forces fxx, fyy, fzz;
!   stats%engcpe, vircpe, engsrp, virsrp and stress(1:9); vdws%elrc, vlrc;
!   n_vdw, max_grid, the number of type-pair keys and vdws%list(key), then for
!   each potential k: tab_potential(0:max_grid, k) and tab_force(0:max_grid, k)
!   (vdw.F90:169-173; present only when the tables exist).
! The stream file is DLPX_DUMP in the working directory, little-endian.
Module dlpx_ref_dump
  Use kinds,         Only: wp
  Use configuration, Only: configuration_type
  Use site,          Only: site_type
  Use statistics,    Only: stats_type
  Use vdw,           Only: vdw_type
  Implicit None
  Private
  Public :: dlpx_dump_state
Contains
  Subroutine dlpx_dump_state(stats, sites, vdws, config)
    Type(stats_type),         Intent(In) :: stats
    Type(site_type),          Intent(In) :: sites
    Type(vdw_type),           Intent(In) :: vdws
    Type(configuration_type), Intent(In) :: config
    Integer :: u, i, k, nkey, have_tab

    Open (newunit=u, file='DLPX_DUMP', access='stream', form='unformatted', status='replace')
    Write (u) Int(config%natms, 4)
    Do i = 1, config%natms
      Write (u) Int(config%ltg(i), 4), config%parts(i)%fxx, config%parts(i)%fyy, config%parts(i)%fzz
    End Do
    Write (u) stats%engcpe, stats%vircpe, stats%engsrp, stats%virsrp, stats%stress(1:9), vdws%elrc, vdws%vlrc
    nkey = sites%ntype_atom * (sites%ntype_atom + 1) / 2
    have_tab = 0
    If (vdws%n_vdw > 0 .and. Allocated(vdws%tab_potential) .and. Allocated(vdws%list)) have_tab = 1
    Write (u) Int(vdws%n_vdw, 4), Int(vdws%max_grid, 4), Int(nkey, 4), Int(have_tab, 4)
    If (have_tab == 1) Then
      Write (u) (Int(vdws%list(k), 4), k = 1, nkey)
      Do k = 1, vdws%n_vdw
        Write (u) vdws%tab_potential(0:vdws%max_grid, k), vdws%tab_force(0:vdws%max_grid, k)
      End Do
    End If
    Close (u)
  End Subroutine dlpx_dump_state
End Module dlpx_ref_dump
