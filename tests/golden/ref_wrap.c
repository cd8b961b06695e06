/* Test infrastructure only (see ref_dump.F90).  The reference binary is
 * linked with --wrap on two_body_forces (two_body.F90:76; flang's symbol for
 * the module procedure), so its call from calculate_forces (drivers.F90:786)
 * comes here: the real routine runs unchanged, and after its first call the
 * state is dumped.  The 21 arguments are scalars (integers, logicals and
 * derived types), which Fortran passes by address with no hidden arguments,
 * so they pass through as pointers. */
#define A void *
extern void __real__QMtwo_bodyPtwo_body_forces(A, A, A, A, A, A, A, A, A, A, A, A, A, A, A, A, A, A, A, A, A);
extern void _QMdlpx_ref_dumpPdlpx_dump_state(A stats, A sites, A vdws, A config);

void __wrap__QMtwo_bodyPtwo_body_forces(A ensemble, A lbook, A megfrz, A leql, A nsteql, A nstep, A stats, A ewld,
                                        A met, A pois, A neigh, A sites, A vdws, A rdf, A mpoles, A electro,
                                        A domain, A tmr, A kim_data, A config, A comm)
{
    static int calls = 0;
    __real__QMtwo_bodyPtwo_body_forces(ensemble, lbook, megfrz, leql, nsteql, nstep, stats, ewld, met, pois, neigh,
                                       sites, vdws, rdf, mpoles, electro, domain, tmr, kim_data, config, comm);
    if (calls++ == 0) _QMdlpx_ref_dumpPdlpx_dump_state(stats, sites, vdws, config);
}
