/*
 * dlpx.h -- C ABI of the MI355X force path for DL_POLY-style molecular dynamics.
 *
 * The library replaces DL_POLY 5.1's per-step force evaluation on one GPU:
 * the link-cell Verlet neighbour list, the real-space two-body loop
 * (tabulated vdW + erfc-damped Coulomb) and the SPME reciprocal sum.  A
 * Fortran host reaches it through ISO_C_BINDING (plain pointers, sizes and
 * byte strides; no C++ types cross the boundary).
 *
 * Conventions (all in DL_POLY internal units: Angstrom, ps, Dalton, 10 J/mol, e):
 *   - cell[9] holds the cell vectors a, b, c as rows (configuration_type%cell).
 *     Orthorhombic cells only (imcon 1 and 2); others return DLPX_ERR_UNSUPPORTED.
 *   - Positions/forces are addressed by base pointer + byte stride, so both
 *     DL_POLY's corePart array (stride 64: xxx,yyy,zzz,fxx,fyy,fzz,chge) and
 *     plain SoA arrays (stride 8) are accepted without copies on the host.
 *   - ltype is DL_POLY's 1-based atom type (config%ltype).
 *   - vdW pair keys are 0-based: key = hi*(hi-1)/2 + lo - 1 for 1-based types
 *     (vdw.F90:1875-1881).
 *   - Every entry point returns 0 on success or a DL_POLY error code
 *     (106 list overflow, 95 cutoff vs cell width, 307 link cells, 120 singular
 *     cell) or a DLPX_ERR_* code; dlpx_last_error() gives the message.
 *   - A context is used from one host thread at a time; every call is
 *     synchronous on return.
 */
#ifndef DLPX_H
#define DLPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DLPX_OK 0
#define DLPX_ERR_LIST_OVERFLOW 106  /* neighbours.F90:1189-1194 */
#define DLPX_ERR_CUTOFF 95          /* neighbours.F90:409-412 */
#define DLPX_ERR_LINK_CELLS 307     /* neighbours.F90:431 */
#define DLPX_ERR_SINGULAR_CELL 120  /* ewald_spole.F90:346 */
#define DLPX_ERR_BAD_ARG 1001
#define DLPX_ERR_UNSUPPORTED 1002
#define DLPX_ERR_STATE 1003
#define DLPX_ERR_HIP 2000           /* + hipError_t */

#define DLPX_FP64 0                 /* FP64 everywhere */
#define DLPX_MIXED 1                /* FP32 pair/erfc, spreading, FFT, gather; FP64 accumulation */

typedef struct dlpx_params {
    int32_t precision;      /* DLPX_FP64 or DLPX_MIXED */
    int32_t natms_max;      /* capacity in atoms (config%mxatms), below 2^25 (else DLPX_ERR_UNSUPPORTED) */
    int32_t ntypes;         /* number of atom types (sites%ntype_atom) */
    int32_t max_list;       /* neighbours per atom (neigh%max_list); 0 = grow as needed */
    double cutoff;          /* neigh%cutoff (real-space Coulomb and list) */
    double padding;         /* neigh%padding (Verlet shell) */
    double vdw_cutoff;      /* vdws%cutoff */
    int32_t electrostatics; /* 0 = none, 1 = SPME (electro%key == ELECTROSTATIC_SPME) */
    int32_t nsplines;       /* ewld%bspline%num_splines */
    int32_t kvec[3];        /* ewld%kspace%k_vec_dim, each 2^a 3^b 5^c */
    double alpha;           /* ewld%alpha */
    double epsilon;         /* electro%eps */
} dlpx_params;

typedef struct dlpx_atoms {
    int32_t natms;
    int32_t stride;         /* bytes between consecutive atoms in every array below */
    const double *x, *y, *z;
    const double *chge;
    double *fx, *fy, *fz;   /* forces are ADDED (callers zero them first) */
    const int32_t *ltype;   /* 1-based, stride 4 bytes */
} dlpx_atoms;

/* Results of one two_body_forces call. */
typedef struct dlpx_two_body_out {
    double buffer[20];      /* the gsum buffer of two_body.F90:708-727 */
    double stress[9];       /* added to stats%stress (row-major 3x3) */
    double engcpe, vircpe;  /* increments of stats%engcpe/vircpe (two_body.F90:766-767) */
    double engsrp, virsrp;  /* increments of stats%engsrp/virsrp incl. vdW LRC (:780-781) */
} dlpx_two_body_out;

typedef struct dlpx_ctx dlpx_ctx;

/* Context creation (ewald_spme_init / init_list / set_bounds sizing). */
int dlpx_create(const dlpx_params *params, dlpx_ctx **ctx);
int dlpx_destroy(dlpx_ctx *ctx);
const char *dlpx_last_error(const dlpx_ctx *ctx);
int dlpx_version(void);

/* Tables.  dlpx_set_vdw_tables takes the host's vdws%tab_potential and
 * vdws%tab_force (vdw.F90:1397-1576): nkeys rows of max_grid+1 values
 * (index 0..max_grid), row = pair key.  ltp[key] is the potential kind
 * (vdw.F90:64-113 keys, -1 = no interaction).  dlpx_set_erfc_tables takes
 * electro%erfc%table and electro%erfc_deriv%table (nsamples values each,
 * value i-1 at r = i*rcut/(nsamples-4); electrostatic.F90:88-127). */
int dlpx_set_vdw_tables(dlpx_ctx *ctx, int32_t nkeys, int32_t max_grid, const double *tab_potential,
                        const double *tab_force, const int32_t *ltp);
int dlpx_set_erfc_tables(dlpx_ctx *ctx, int32_t nsamples, const double *erfc, const double *erfc_deriv);
/* The same tables generated in the library from potential parameters
 * (vdw_generate, vdw.F90:1397-1576): kind[key] is the vdw.F90:64-113 key of
 * each pair key, -1 none; params holds vdws%param(1:7, key), 7 doubles per
 * key in the reference's own layout (bounds.F90:666 max_param = 7), in the
 * internal units the FIELD reader leaves (ffield.F90:3757-3839; LJF's
 * converted ea, sig^2, rc^2).  Forms (two_body_potentials.F90): 1 = 12-6 A,B;
 * 2 = Lennard-Jones eps,sig; 3 = n-m e0,n,m,r0; 4 = Buckingham A,rho,C;
 * 5 = Born-Huggins-Meyer A,B,sigma,C,D; 6 = 12-10 A,B; 7 = shifted n-m
 * e0,n,m,r0,r_trunc; 8 = Morse e0,r0,k; 9 = WCA eps,sig,d,cut (cut =
 * 2^(1/6) sig + d, ffield.F90:3788); 10 = DPD a,rc; 11 = AMOEBA 14-7 eps,sig;
 * 12 = LJ cohesive eps,sig,c; 13 = Morse + c/r^12 e0,r0,k,c; 14 = Rydberg
 * a,b,c; 15 = ZBL Z1,Z2; 16 = ZBL switched to Morse Z1,Z2,rm,ic,e0,r0,k;
 * 17 = ZBL switched to Buckingham Z1,Z2,rm,ic,A,rho,C; 18 = LJ x MDF
 * eps,sig,ri; 19 = Buckingham x MDF A,rho,C,ri; 20 = 12-6 x MDF a,b,ri (the
 * damping ends at the vdW cutoff, vdw.F90:205-230); 21 = LJF ea,sig^2,rc^2;
 * 22 = Sanderson A,L,d; 23 = nDPD a,b,n,rc; 24 = Stillinger-Weber two-body
 * eps,A,B,sig,p,q,aa.  Tabulated (TABLE file) potentials: use
 * dlpx_set_vdw_tables.  Sizes follow bounds.F90:811,820. */
int dlpx_generate_vdw_tables(dlpx_ctx *ctx, int32_t nkeys, const int32_t *kind, const double *params);
int dlpx_generate_erfc_tables(dlpx_ctx *ctx);
/* vdW long-range corrections vdws%elrc / vdws%vlrc (vdw.F90:617-967).  Pass the
 * host's values, or (after dlpx_generate_vdw_tables) call dlpx_compute_lrc
 * with the number of atoms of each type. */
int dlpx_set_lrc(dlpx_ctx *ctx, double elrc, double vlrc);
int dlpx_compute_lrc(dlpx_ctx *ctx, const double cell[9], const double *num_type);
/* vdW force-shifting (vdws%l_force_shift, CONTROL vdw_force_shift; vdw.F90:
 * 1922,1961-1964): each pair's vdW gamma loses tab_force(rvdw)/(r rvdw) and its
 * energy gains tab_force(rvdw) (r/rvdw - 1) - tab_potential(rvdw), with the
 * table samples at r = rvdw (index max_grid-4).  The reference then applies no
 * long-range correction (vdw.F90:662): switching it on zeroes elrc and vlrc,
 * and dlpx_compute_lrc leaves them zero.  on: 0 or 1. */
int dlpx_set_vdw_force_shift(dlpx_ctx *ctx, int32_t on);
/* vdw_method direct (vdws%l_direct; vdw.F90:1590-1788 vdw_forces_direct): each
 * pair evaluates its analytic form (the forms of dlpx_generate_vdw_tables,
 * which must have been called) instead of interpolating the tables; with
 * force-shifting, U gains afs r + bfs and gamma loses afs/r, afs =
 * gamma(rvdw)/rvdw, bfs = -U(rvdw) - gamma(rvdw) (vdw.F90:969-1048).  on: 0 or 1. */
int dlpx_set_vdw_direct(dlpx_ctx *ctx, int32_t on);
int dlpx_get_lrc(dlpx_ctx *ctx, double *elrc, double *vlrc);

/* Excluded pairs, neigh%list_excl as build_excl_intra fills it from bonds,
 * angles, dihedrals, constraints, rigid bodies and core-shell units
 * (build_excl.F90:32-1195).  list_excl holds mxexcl+1 int32 per atom, laid
 * out as the reference's list_excl(0:mxexcl, 1:natms): element k of atom i
 * (0-based host index) at list_excl[(mxexcl+1)*i + k], k = 0 the count, then
 * the partners' 1-based global ids (ltg).  Each pair must be listed from both
 * ends.  Excluded pairs leave the Verlet list (neighbours.F90:1229-1250) and,
 * with SPME, get the Ewald exclusion correction U = -q_i q_j erf(alpha r)/r
 * for r < cutoff (ewald_spole.F90:479-679), reported in buffer slots 15-16
 * (two_body.F90:723-724) and included in engcpe, vircpe and the stress.
 * Call before dlpx_link_cell_pairs / dlpx_md_load; natms = 0 clears them. */
int dlpx_set_exclusions(dlpx_ctx *ctx, int32_t natms, int32_t mxexcl, const int32_t *list_excl);

/* Frozen atoms, config%lfrzn (nonzero = frozen; 0-based host order, natms
 * values).  A pair of two frozen atoms leaves the Verlet list
 * (neighbours.F90:1198-1225), so it has no vdW or real-space Coulomb term;
 * the reciprocal sum keeps it (DL_POLY 5.1 calls no frozen-pair Ewald
 * correction).  In resident runs (dlpx_md_load / dlpx_md_run) frozen atoms
 * get zero force and velocity after each force evaluation, as freeze_atoms
 * does (configuration.F90:3240-3258); dlpx_two_body_forces leaves that to
 * the host.  Call before dlpx_link_cell_pairs / dlpx_md_load; natms = 0
 * clears them. */
int dlpx_set_frozen(dlpx_ctx *ctx, int32_t natms, const int32_t *lfrzn);
/* dlpx_compute_lrc with num_frozen[type] frozen atoms of each type: the
 * frozen-frozen pairs carry no correction (vdw.F90:665-670,953). */
int dlpx_compute_lrc_frozen(dlpx_ctx *ctx, const double cell[9], const double *num_type, const double *num_frozen);

/* Per-particle energies and stresses, stats%collect_pp (statistics.F90:368-
 * 402; used for HEATFLUX and per-particle output, drivers.F90:2117-2165).
 * dlpx_set_per_particle(ctx, 1) makes each following dlpx_two_body_forces
 * collect them as the reference does: half of every pair's energy and of
 * r (x) f to each atom (vdw.F90:1750-1760,1987-1997, ewald_spole.F90:205-
 * 213; a tabulated-vdW pair across the cell boundary gives its energy share
 * only to the atom with the smaller global id, vdw.F90:1946-1966), the atom's
 * reciprocal energy q sum M Phi and its stress from the stress-kernel grid
 * (ewald_general.F90:578-715,871-1009,1033-1057), and the self energy and
 * vdW long-range correction shared evenly (ewald_spole.F90:465,
 * two_body.F90:783).  As in the reference, the reciprocal stress of those
 * calls (and so vircpe and the stress output) is then the sum of the
 * per-particle ones.  Needs the default 8 pair lanes.
 * dlpx_add_per_particle adds the last call's values into pp_energy(natms)
 * and pp_stress(9, natms) (column-major 3x3 per atom, element a+3b = r_a f_b
 * summed); either pointer may be NULL. */
int dlpx_set_per_particle(dlpx_ctx *ctx, int32_t on);
int dlpx_add_per_particle(dlpx_ctx *ctx, double *pp_energy, double *pp_stress);

/* link_cell_pairs (neighbours.F90:356-1414): build the Verlet list of pairs
 * within cutoff+padding from the given positions.  On success *npairs is the
 * number of unordered pairs and *ibig the largest per-atom count; with a fixed
 * max_list a larger count returns 106 with *ibig set. */
int dlpx_link_cell_pairs(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms, int64_t *npairs,
                         int32_t *ibig);
/* The current list as unordered pairs (i < j, 0-based host indices) with the
 * image shift s such that r_ij = x_i - (x_j + s.L); 5 int32 per pair.
 * *npairs returns the count; at most cap pairs are written. */
int dlpx_get_pairs(dlpx_ctx *ctx, int32_t *pairs, int64_t cap, int64_t *npairs);
/* Inspection, no DL_POLY counterpart: the number of unordered pairs of the
 * current list closer than r at the device's current positions (FP64 minimum
 * image), counted on the device; for a throughput figure in pairs within the
 * cutoff (the list also holds the padding shell). */
int dlpx_count_pairs(dlpx_ctx *ctx, double r, int64_t *npairs);

/* vnl_check (neighbours.F90:123-296): *update = 1 when some atom moved by
 * at least padding/2 since the list was built. */
int dlpx_vnl_check(dlpx_ctx *ctx, const dlpx_atoms *atoms, int32_t *update);

/* two_body_forces (two_body.F90:76-796) for vdW + SPME Coulomb on the current
 * list: forces are added into atoms->fx/fy/fz; energies, virials and stress as
 * in dlpx_two_body_out. */
int dlpx_two_body_forces(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms,
                         dlpx_two_body_out *out);
/* dlpx_two_body_forces without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in DL_POLY.  Same inputs, outputs and
 * results; the partial sums live in a fixed scratch array of the library
 * (7.3 MB per device: up to about a million atoms at spline order 8; larger
 * systems return DLPX_ERR_BAD_ARG).  Preconditions: the list was built with
 * dlpx_link_cell_pairs after the last table change (DLPX_ERR_STATE
 * otherwise), and per-particle output is off (DLPX_ERR_UNSUPPORTED).  Every
 * context on a device shares that device's scratch, so calls that overlap in
 * time on one device (from several host threads) race: serialise them. */
int dlpx_two_body_forces_noalloc(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms,
                                 dlpx_two_body_out *out);

/* Resident mode: positions, velocities and forces stay in HBM between steps.
 * dlpx_md_load copies a configuration in: vx/vy/vz and mass are read with
 * byte stride vstride (0 = atoms->stride), so DL_POLY's config%vxx/vyy/vzz and
 * config%weight (stride 8) go in beside corePart positions (stride 64).
 * dlpx_md_run performs nsteps velocity-Verlet NVE steps
 * (nve.F90:127-215) with the displacement-triggered list rebuild and the
 * two-body forces of every step; when energies is not NULL, step s writes
 * energies[5*s + {0..4}] = engcpe, vircpe, engsrp, virsrp (as in
 * dlpx_two_body_out) and the kinetic energy.  dlpx_md_store copies positions,
 * velocities and forces back (any pointer may be NULL): positions and forces
 * with byte stride stride, velocities with vstride (0 = stride). */
int dlpx_md_load(dlpx_ctx *ctx, const double cell[9], const dlpx_atoms *atoms, const double *vx,
                 const double *vy, const double *vz, const double *mass, int32_t vstride);
int dlpx_md_run(dlpx_ctx *ctx, int32_t nsteps, double tstep, double *energies, int32_t *nrebuilds);
int dlpx_md_store(dlpx_ctx *ctx, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                  double *fx, double *fy, double *fz, int32_t stride, int32_t vstride);

/* The two-body output (energies, virials, stress and the 20-slot buffer, as
 * dlpx_two_body_forces returns them) of the last step of the most recent
 * dlpx_md_run call that recorded energies, or of the starting configuration
 * after dlpx_md_load: what two_body_forces added into stats at that step
 * (two_body.F90:766-790), for a host that writes the step's statistics
 * (statistics.F90:1118-1160) without copying the configuration back.
 * DLPX_ERR_STATE before any dlpx_md_load. */
int dlpx_md_last_output(dlpx_ctx *ctx, dlpx_two_body_out *out);

/* Batched ensemble: the same as dlpx_md_run(ctxs[k], nsteps, tsteps[k], NULL,
 * &nrebuilds[k]) for every k, with every stage of a step one launch over all
 * the systems: per-system argument tables in device memory, each workgroup
 * finding its system; the list rebuild of each system decided on the device
 * by its own stage 1; the steps replaying one captured graph, with no host
 * round trip until the run ends.  Launch shapes (pair-kernel lanes per atom,
 * FFT lines per workgroup) are chosen per size class by timing candidates on
 * the first steps of a long enough run, and cached (DLPX_AUTOTUNE=0 disables
 * that; DLPX_TUNE_CACHE names a cache file).  Every context must have been
 * loaded with dlpx_md_load, be on the same device, appear once and share the
 * precision mode; systems with electrostatics must share the spline order.
 * The contexts' work runs on the first context's streams.  No counterpart in
 * DL_POLY, which runs one system per process (calculate_forces,
 * drivers.F90:573-929): it serves hosts that run many small independent
 * systems.  dlpx_md_run_ensemble_energies also records each system's per-step
 * energies as dlpx_md_run does, system k's step s at energies[5*(nsteps*k +
 * s) + {0..4}] (energies may be NULL). */
int dlpx_md_run_ensemble(dlpx_ctx *const *ctxs, int32_t nctx, int32_t nsteps, const double *tsteps,
                         int32_t *nrebuilds);
int dlpx_md_run_ensemble_energies(dlpx_ctx *const *ctxs, int32_t nctx, int32_t nsteps, const double *tsteps,
                                  double *energies, int32_t *nrebuilds);

/* Per-stage device time (ms, HIP events on the context's stream), summed over
 * the force evaluations since dlpx_set_timing(ctx, 1): ms[0] pair forces,
 * ms[1] B-spline charge spreading, ms[2] FFT + influence function + inverse
 * FFT, ms[3] force gather, ms[4] reductions and reciprocal-force finish,
 * ms[5] neighbour-list kernel, ms[7] number of evaluations timed. */
int dlpx_set_timing(dlpx_ctx *ctx, int32_t enable);
int dlpx_get_timing(dlpx_ctx *ctx, double *ms, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
