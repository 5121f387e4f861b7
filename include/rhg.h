/*
 * rhg.h -- C ABI of the MI355X host-guest energy engine (librhg.so).
 *
 * Drop-in boundary for RASPA2's host-guest energy path: framework van der
 * Waals and real-space Ewald energies at trial positions, the Fourier-space
 * Ewald change of an inserted molecule, Widom/CBMC insertion, the Widom
 * estimators, and energy grids.  Plain pointers and sizes only; every entry
 * point returns an int status (RHG_OK or a negative RHG_ERR_*) and leaves a
 * message for rhg_last_error().
 *
 * Unit conventions follow RASPA2: positions in Angstrom, charges in e, LJ
 * epsilon in K on input, energies in RASPA2's internal unit
 * (amu*A^2/ps^2 = 1.6605402e-23 J; src/constants.c:147).
 *
 * Threading: calls on one context are serialised by the caller; distinct
 * contexts are independent.  Each context owns one HIP stream.
 */
#ifndef RHG_H
#define RHG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RHG_OK 0
#define RHG_ERR_INVALID (-1)
#define RHG_ERR_HIP (-2)
#define RHG_ERR_OOM (-3)
#define RHG_ERR_NOT_ADMITTED (-4)
#define RHG_ERR_UNSUPPORTED (-5)
#define RHG_ERR_FAULT_INJECTED (-6)
#define RHG_ERR_IO (-7)

/* BoundaryCondition values (src/potentials.h:232-271) */
#define RHG_BC_RECTANGULAR 1
#define RHG_BC_TRICLINIC 2

/* ChargeMethod values used on this path (src/potentials.c:12430-12478) */
#define RHG_CHARGE_NONE 0
#define RHG_CHARGE_EWALD 1

/* arithmetic of the pair kernels */
#define RHG_PREC_FP64 0  /* plain: FP64 throughout, RASPA2's operation order */
#define RHG_PREC_MIXED 1 /* FP32 pair terms, FP64 differences and accumulation */

#define RHG_MAX_SITES 8
#define RHG_NR_BLOCKS 5 /* src/statistics.h:40 */

typedef struct rhg_ctx rhg_ctx;
typedef struct rhg_workspace rhg_workspace;
typedef struct rhg_grid rhg_grid;

/* Framework atoms, one pseudo-atom per atom (Lorentz-Berthelot mixing on
 * the fly).  Replaces Framework[s].Atoms[f][i] (src/framework.h:180,
 * src/molecule.h:227-263) and the framework rows of PseudoAtoms[]. */
typedef struct {
  int32_t n_atoms;
  const double *x, *y, *z; /* Cartesian positions [A] */
  const double *charge;    /* [e]; NULL = uncharged framework */
  const double *eps_k;     /* LJ epsilon/k_B [K] */
  const double *sigma;     /* LJ sigma [A] */
  const double *mass;      /* [g/mol]; NULL = 12.0 each */
} rhg_framework;

/* Simulation box: RASPA2's Box matrix (fields ax ay az bx by bz cx cy cz,
 * the columns are the cell vectors; src/framework.c:3049-3066). */
typedef struct {
  double box[9];
  int32_t boundary; /* RHG_BC_* */
} rhg_cell;

typedef struct {
  double cutoff_vdw;      /* CutOffVDW [A] */
  int32_t shift_vdw;      /* 1: LJ shifted to zero at the cutoff */
  int32_t charge_method;  /* RHG_CHARGE_* */
  double cutoff_charge;   /* CutOffChargeCharge [A] */
  double ewald_precision; /* EwaldPrecision (automatic alpha and kmax) */
  double temperature;     /* ExternalTemperature [K] */
} rhg_forcefield;

/* A rigid adsorbate (one rigid group holding all atoms). */
typedef struct {
  int32_t n_sites;
  int32_t type[RHG_MAX_SITES];        /* pseudo-atom ids from rhg_add_pseudo_atom */
  double body[RHG_MAX_SITES][3];      /* Components[c].Positions [A] */
  int32_t start;                      /* StartingBead */
  int32_t k_fb;                       /* NumberOfTrialPositionsForTheFirstBeadWidom */
  int32_t k_or;                       /* NumberOfTrialPositionsWidom */
} rhg_molecule;

/* Widom block accumulators (WidomRosenbluthFactorCount/Accumulated,
 * WidomEnergyDifferenceAccumulated, WidomIdealGasAccumulated;
 * src/mc_moves.c:7862-7933). */
typedef struct {
  double count[RHG_NR_BLOCKS];
  double sum_w[RHG_NR_BLOCKS];
  double sum_wdu[RHG_NR_BLOCKS];
  double ideal_gas[RHG_NR_BLOCKS];
} rhg_widom_blocks;

/* per-insertion record layout (rec arrays are n x RHG_REC_N doubles) */
#define RHG_REC_W 0        /* RosenbluthNew incl. Fourier factor; 0 on overlap */
#define RHG_REC_DU 1       /* DeltaU of WidomAdsorbateMove */
#define RHG_REC_UVDW 2     /* UHostVDWNew */
#define RHG_REC_UREAL 3    /* UHostChargeChargeNew (real space) */
#define RHG_REC_UFHOST 4   /* UHostAdsorbateChargeChargeFourierDelta */
#define RHG_REC_UFADS 5    /* UAdsorbateAdsorbateChargeChargeFourierDelta */
#define RHG_REC_OVERLAP 6  /* 1: rejected (OVERLAP) */
#define RHG_REC_SELFB 7
#define RHG_REC_SELOR 8
#define RHG_REC_FBMASK 9   /* overlapping first-bead trials, bit i */
#define RHG_REC_ORMASK 10  /* overlapping orientation trials, bit i */
#define RHG_REC_W1 11
#define RHG_REC_W2 12
#define RHG_REC_STAGE 13   /* 0 fb all overlap,1 W1<min,2 or all overlap,3 W<min,4 accepted,5 blocked */
#define RHG_REC_UADSVDW 14 /* UAdsorbateVDWNew (guest-guest, finite loading) */
#define RHG_REC_UADSREAL 15 /* UAdsorbateChargeChargeNew (real space) */
#define RHG_REC_N 16

/* ------------------------------------------------------------ context */
/* Replaces ReadInput's framework/force-field/Ewald set-up for one system:
 * ReadFrameworkDefinition, mixing (src/potentials.c:3164-3171),
 * ComputePotentialShifts (src/potentials.c:5263), InitializeEwald
 * (src/ewald.c:200), SetupKVectors (src/ewald.c:532), EwaldEnergyIon
 * (src/ewald.c:615) and PrecomputeTotalEwaldContributions
 * (src/ewald.c:1289).  device = HIP device ordinal. */
int rhg_create(const rhg_framework *fw, const rhg_cell *cell, const rhg_forcefield *ff, int device,
               rhg_ctx **out);
void rhg_destroy(rhg_ctx *ctx);
const char *rhg_last_error(const rhg_ctx *ctx);
const char *rhg_version(void);
/* Adds a probe pseudo-atom (a PseudoAtoms[] row, src/potentials.c:268-274). */
int rhg_add_pseudo_atom(rhg_ctx *ctx, double eps_k, double sigma, double charge, int32_t *type_out);
int rhg_set_precision(rhg_ctx *ctx, int32_t precision);
/* HBM bytes held by the context (framework tables, k tables) */
int rhg_footprint(const rhg_ctx *ctx, size_t *bytes);
/* out: alpha, kmax_x, kmax_y, kmax_z, nk, uion, beta, volume, framework mass, density [kg/m^3] */
int rhg_info(const rhg_ctx *ctx, double out[10]);

/* ------------------------------------------------- energy at a position */
/* REAL CalculateFrameworkVDWEnergyAtPosition(POINT,int,REAL)
 * (src/framework_energy.h:43, src/framework_energy.c:60) */
int rhg_framework_vdw_energy_at_position(rhg_ctx *ctx, const double pos[3], int32_t type,
                                         double scaling, double *u_vdw);
/* void CalculateFrameworkChargeEnergyAtPosition(POINT,int,REAL*,REAL*,REAL)
 * (src/framework_energy.h:46, src/framework_energy.c:176); ucbd is 0
 * (no bond dipoles on this path). */
int rhg_framework_charge_energy_at_position(rhg_ctx *ctx, const double pos[3], int32_t type,
                                            double scaling, double *ucc, double *ucbd);
/* Batched form of both: n positions (xyz interleaved) of one pseudo-atom
 * type.  overlap[i] = (u_vdw >= 1e7 || u_cc >= 1e7), EnergyOverlapCriteria
 * of src/cbmc.c:699-712.  Any output may be NULL. */
int rhg_framework_energy_at_positions(rhg_ctx *ctx, int64_t n, const double *xyz, int32_t type,
                                      double *u_vdw, double *u_cc, uint8_t *overlap);
/* No counterpart in the reference.  rhg_framework_energy_at_positions for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  It calls no
 * hipMalloc/hipFree and no allocating library routine.  Positions and
 * outputs are DEVICE pointers (d_xyz: n x 3 interleaved; outputs nullable),
 * the results equal the allocating form's, and the work is only enqueued on
 * the context's stream (rhg_synchronize waits for it).  Binning scratch is a
 * fixed device array of the library (18 MiB, one per GPU, held until the
 * process ends): n <= 262144 (RHG_ERR_INVALID beyond).  The context's cell
 * lists must exist (RHG_ERR_INVALID otherwise): one call of the allocating
 * form at set-up builds them.  Every call on one GPU shares that GPU's array,
 * so calls on different contexts (streams) of one GPU that overlap in time
 * race; calls on one context are ordered by its stream. */
int rhg_framework_energy_at_positions_noalloc(rhg_ctx *ctx, int64_t n, const double *d_xyz, int32_t type,
                                              double *d_u_vdw, double *d_u_cc, uint8_t *d_overlap);

/* ---------------------------------------------------------------- Ewald */
/* Ewald tables in RASPA2's k order (kx 0..kmax, ky -kmax..kmax,
 * kz -kmax..kmax): KFactor and StoreTotalChargeFramework
 * (src/ewald.c:532-600, 1289-1731).  Arrays hold nk values. */
int rhg_ewald_tables(rhg_ctx *ctx, double *kfactor, double *sf_re, double *sf_im);
/* int CalculateEwaldFourierAdsorbate(int New,int Old,int mol,int store)
 * (src/ewald.h:86, src/ewald.c:6859) for n_mol independent new molecules
 * at zero loading: d_host = UHostAdsorbateChargeChargeFourierDelta,
 * d_ads = UAdsorbateAdsorbateChargeChargeFourierDelta.  xyz: n_mol x
 * n_sites x 3. */
int rhg_ewald_fourier_adsorbate(rhg_ctx *ctx, int64_t n_mol, int32_t n_sites, const int32_t *types,
                                const double *xyz, double *d_host, double *d_ads);

/* ---------------------------------------------------------------- Widom */
/* WidomAdsorbateMove (src/mc_moves.c:7844) for insertions [first, first+n)
 * of a job of n_total insertions and steps_per_cycle MC steps per cycle
 * (the block of insertion i is that of cycle i/steps_per_cycle,
 * src/statistics.c:904,1126).  Trial positions come from the per-insertion
 * counter stream of `seed` in RASPA2's per-move draw order (first-bead z,y,x
 * src/cbmc.c:652-667; Marsaglia sphere and Shoemake rotation
 * src/utils.c:688,835).  Blocks are ADDED to *blocks.  rec (nullable)
 * receives n x RHG_REC_N doubles. */
int rhg_widom_run(rhg_ctx *ctx, const rhg_molecule *mol, uint64_t seed, int64_t n_total,
                  int64_t first, int64_t n, int32_t steps_per_cycle, rhg_widom_blocks *blocks,
                  double *rec);
/* Same move with host-supplied trials (the caller's own RNG, e.g. RASPA2's
 * MT19937-64): fb n x k_fb x 3 first-bead positions, cord n x k_or x
 * n_sites x 3 rotated offsets from the starting bead, u_sel n x 2 selection
 * uniforms.  No blocks; rec is required. */
int rhg_widom_run_trials(rhg_ctx *ctx, const rhg_molecule *mol, int64_t n, const double *fb,
                         const double *cord, const double *u_sel, double *rec);
/* Blocked pockets of the Widom path (src/framework.c:6168-6300, checked at
 * src/mc_moves.c:7882-7886): n centres (xyz, n x 3, already replicated over
 * the unit cells of the system as src/framework.c:6209-6226 does) with radii
 * [A], periodic with the unit cell unit_box (ApplyBoundaryConditionUnitCell,
 * src/potentials.c:5715).  A grown molecule with any atom within a pocket
 * (outside every pocket when invert = 1, InvertBlockPockets) contributes
 * W = 0 with record stage 5.  Applies to later Widom calls on the context;
 * n = 0 removes the pockets. */
int rhg_set_block_pockets(rhg_ctx *ctx, int32_t n, const double *xyz, const double *radius,
                          const double unit_box[9], int32_t invert);
/* int BlockedPocket(VECTOR pos) (src/framework.c:6246) for n positions */
int rhg_blocked_pocket(rhg_ctx *ctx, int64_t n, const double *xyz, uint8_t *blocked);
/* Asynchronous form of rhg_widom_run, for callers that keep several systems
 * in flight (RASPA2 runs its systems one after another, src/monte_carlo.c:
 * 626-748; a screening loop over many frameworks wants them concurrent):
 * rhg_widom_submit enqueues the insertions on the context's stream and
 * returns; the block sums of every submit since the last collect accumulate
 * on the device.  rhg_widom_query returns 1 when the stream is idle, 0 if
 * work is pending.  rhg_widom_collect waits, ADDS the sums to *blocks
 * (nullable) and clears them.  Results are bitwise those of rhg_widom_run
 * over the same ranges. */
int rhg_widom_submit(rhg_ctx *ctx, const rhg_molecule *mol, uint64_t seed, int64_t n_total,
                     int64_t first, int64_t n, int32_t steps_per_cycle);
int rhg_widom_query(rhg_ctx *ctx);
int rhg_widom_collect(rhg_ctx *ctx, rhg_widom_blocks *blocks);
/* Execution options of the Widom path (no counterpart in the reference):
 * chunk = insertions per pipeline pass (1..2^24, default 2^20; the block
 * sums are grouped by chunk, so equal chunks give bitwise equal sums);
 * graphs = 1 replays chunks after the first as one captured HIP graph. */
int rhg_set_widom_options(rhg_ctx *ctx, int64_t chunk, int32_t graphs);
/* Host-side fault hook for failure-handling tests (no counterpart in the
 * reference): the chunk after `after_chunks` more chunks returns
 * RHG_ERR_FAULT_INJECTED before launching anything; < 0 disables it. */
int rhg_inject_fault(rhg_ctx *ctx, int64_t after_chunks);
/* GetAverageHenryCoefficient / block prints (src/statistics.c:2539,
 * 4747-4800): out[0..4] K_H per block [mol/kg/Pa], out[5] average, out[6]
 * 2 sigma, out[7..11] <U_gh>_1-<U_h>_0 per block [K], out[12] average,
 * out[13] 2 sigma, out[14] overall K_H, out[15] overall <W>. */
int rhg_widom_estimators(const rhg_widom_blocks *blocks, double temperature, double density_kg_m3,
                         double ideal_gas_weight, double out[16]);

/* -------------------------------------------------------- finite loading */
/* The adsorbate molecules present (Adsorbates[s], finite loading): molecule
 * m holds atoms [start[m], start[m+1]) of the given pseudo-atom types at xyz
 * (n x 3 doubles, n = start[n_mol]); n_mol = 0 clears.  The adsorbate
 * structure factor (StoreTotalChargeAdsorbates) and net charge are computed
 * on the device (src/ewald.c:1289-1731).  From then on Widom insertion adds
 * the guest-guest terms and the Fourier change uses the stored sums. */
int rhg_set_adsorbates(rhg_ctx *ctx, int32_t n_mol, const int32_t *start, const int32_t *types,
                       const double *xyz);
/* REAL CalculateInterVDWEnergyAdsorbateAtPosition(POINT,int,int,REAL) and
 * int CalculateInterChargeEnergyAdsorbateAtPosition(POINT,int,REAL*,REAL*,int,REAL)
 * (src/inter_energy.h, src/inter_energy.c:321-520) at n positions for
 * pseudo-atom `type` (charge scaling 1), every molecule but `exclude`
 * (-1: none).  u_vdw, u_cc nullable. */
int rhg_inter_energy_at_positions(rhg_ctx *ctx, int64_t n, const double *xyz, int32_t type, int32_t exclude,
                                  double *u_vdw, double *u_cc);
/* REAL RetraceMolecule(CBMC_DELETION) (src/cbmc.h, src/cbmc.c:2928-3054) for
 * n existing molecules: old_xyz holds each one's n_sites x 3 positions; fb
 * (n x k_fb x 3) and cord (n x k_or x n_sites x 3) the other trials, whose
 * entries 0 are ignored (the molecule's own configuration takes that slot
 * and is always chosen).  exclude[i] is molecule i's adsorbate index (-1:
 * none; exclude null: none for all) so that it does not interact with
 * itself.  rec as rhg_widom_run (W = RosenbluthOld; no Fourier term, which
 * rhg_ewald_fourier_move gives).  A configuration that itself overlaps is
 * reported as overlap (stage 0 or 2). */
int rhg_widom_retrace_trials(rhg_ctx *ctx, const rhg_molecule *mol, const int32_t *exclude, int64_t n,
                             const double *old_xyz, const double *fb, const double *cord, double *rec);
/* int CalculateEwaldFourierAdsorbate(int New,int Old,int mol,int store)
 * (src/ewald.h:86, src/ewald.c:6859-7565) for one move: the new molecule's
 * n_new sites (0: none) and/or the removal of adsorbate molecule old_mol
 * (-1: none), against the stored framework and adsorbate sums.  out[0]
 * UHostAdsorbateChargeChargeFourierDelta, out[1]
 * UAdsorbateAdsorbateChargeChargeFourierDelta.  At most 8 charged sites in
 * total (RHG_ERR_UNSUPPORTED). */
int rhg_ewald_fourier_move(rhg_ctx *ctx, int32_t n_new, const int32_t *new_types, const double *new_xyz,
                           int32_t old_mol, double out[2]);

/* ---------------------------------------------------------- surface area */
/* SurfaceAreaMove / CheckSurfaceAreaOverlap (src/mc_moves.c:16434-16520,
 * src/framework.c:11649-11720): `samples` points per framework atom on the
 * sphere of radius factor x sigma(probe, atom) (factor 2^(1/6) for RASPA2's
 * default "Minimum", 1 for "Sigma"), a point counting when it lies in no
 * blocked pocket and within that distance of no other atom.  area = sum over
 * atoms of 4 pi d^2 x the counted fraction [A^2]; counts (nullable) the
 * accessible points per atom.  interacts (nullable: every atom) flags the
 * atoms whose pseudo-atom interacts (PseudoAtoms[].Interaction,
 * src/potentials.c:12170-12185): an atom of a NONE pseudo-atom has no sphere
 * (src/mc_moves.c:16456; count 0) and excludes no point (its mixed sigma is
 * 0, src/framework.c:11676).  Points come from the counter stream of
 * (seed, atom x samples + k). */
int rhg_surface_area(rhg_ctx *ctx, int32_t probe_type, double distance_factor, int64_t samples, uint64_t seed,
                     const uint8_t *interacts, double *area, uint64_t *counts);

/* ------------------------------------------------ pore-size distribution */
/* SimulationType PSD (ComputePoreSizeDistribution, src/monte_carlo.c:1515-
 * 1540; SamplePoreSizeDistribution, src/sample.c:3444-3560): `cycles`
 * points A and n_b sphere centres per point (the reference draws
 * 20 x NumberOfCycles), PSDProbeDistance probe_factor (1 for Sigma,
 * 2^(1/6) for Minimum: an atom excludes 0.5 factor sigma_jj), blocked
 * pockets of the context honoured.  hist[bins] (PSDHistogramSize over
 * PSDRange) receives, ADDED, the reference's PoreSizeDistributionHistogram:
 * the number of cycles whose largest sphere holding A reaches bin k or
 * beyond (a cycle beyond the range counts nowhere).  largest (nullable,
 * cycles doubles) receives each cycle's largest radius (-10: none).  Points
 * come from the counter stream of (seed, cycle): A from counters 0..2,
 * centre j from 3+3j..5+3j, each in x, y, z order. */
int rhg_pore_size_distribution(rhg_ctx *ctx, double probe_factor, int64_t cycles, int64_t n_b, uint64_t seed,
                               int32_t bins, double range, uint64_t *hist, double *largest);

/* ------------------------------------------------------- flexible chains */
/* A flexible molecule made of one flexible group (the alkane family of
 * molecules/TraPPE: ReadComponentDefinition, src/molecule.c:1430-3030).
 * Bonded parameters are in the units of the molecule file; the library
 * converts them as the reference does on reading (force constants divided
 * by ENERGY_TO_KELVIN, angles to radians).  Intra-molecular LJ pairs use
 * the pseudo-atoms' mixed parameters; pairs that close a bend-bend quartet
 * get intra14_vdw_scaling (src/molecule.c:2997-3025). */
#define RHG_CHAIN_MAX_ATOMS 64
#define RHG_BOND_FIXED 0     /* FIXED_BOND: p0 = length [A] */
#define RHG_BOND_HARMONIC 1  /* HARMONIC_BOND: 0.5 p0 (r - p1)^2, p0 [K/A^2], p1 [A] */
#define RHG_BEND_HARMONIC 0  /* HARMONIC_BEND: 0.5 p0 (theta - p1)^2, p0 [K/rad^2], p1 [deg] */
#define RHG_TORSION_TRAPPE 0 /* TRAPPE_DIHEDRAL: p0 + p1(1+cos) + p2(1-cos2) + p3(1+cos3) [K] */
typedef struct {
  int32_t n_atoms;              /* 2..RHG_CHAIN_MAX_ATOMS */
  int32_t start;                /* StartingBead */
  const int32_t *type;          /* pseudo-atom of each atom (uncharged) */
  int32_t n_bonds;
  const int32_t *bonds;         /* n_bonds x 3: atom a, atom b, RHG_BOND_* */
  const double *bond_par;       /* n_bonds x 4 */
  int32_t n_bends;
  const int32_t *bends;         /* n_bends x 4: a, b (apex), c, RHG_BEND_* */
  const double *bend_par;       /* n_bends x 4 */
  int32_t n_torsions;
  const int32_t *torsions;      /* n_torsions x 5: a, b, c, d, RHG_TORSION_* */
  const double *torsion_par;    /* n_torsions x 4 */
  int32_t n_intra_vdw;
  const int32_t *intra_vdw;     /* n_intra_vdw x 2 atom pairs */
  double intra14_vdw_scaling;   /* Intra14VDWScalingValue (1.0) */
  int32_t k_fb;                 /* NumberOfTrialPositionsForTheFirstBeadWidom (1..64) */
  int32_t k_trial;              /* NumberOfTrialPositionsWidom (1..64, <= k_torsion) */
  int32_t k_torsion;            /* NumberOfTrialPositionsTorsion (100; 1..1024) */
  int32_t moves_per_open_bead;  /* NumberOfTrialMovesPerOpenBead (150) */
  const double *max_bond_change;   /* per atom MaximumCBMCChangeBondLength [A] (NULL: 0.3) */
  const double *max_bend_change;   /* per atom MaximumCBMCChangeBendAngle [rad] (NULL: 0.3) */
  const double *max_cone_rotation; /* per atom MaximumCBMCRotationOnCone [rad] (NULL: 0.3) */
} rhg_chain;

/* per-insertion chain record layout (n x RHG_CREC_N doubles) */
#define RHG_CREC_W 0        /* RosenbluthNew; 0 when rejected or blocked */
#define RHG_CREC_DU 1       /* DeltaU of WidomAdsorbateMove */
#define RHG_CREC_UHOSTVDW 2 /* UHostVDWNew */
#define RHG_CREC_UINTRA 3   /* UIntraVDWNew */
#define RHG_CREC_UBOND 4    /* UBondNew */
#define RHG_CREC_UBEND 5    /* UBendNew */
#define RHG_CREC_UTORSION 6 /* UTorsionNew */
#define RHG_CREC_OVERLAP 7  /* 1: rejected (OVERLAP) */
#define RHG_CREC_STAGE 8    /* 0 fb all overlap, 1 fb W<min, 2 step all overlap, 3 step W<min, 4 accepted, 5 blocked */
#define RHG_CREC_STEP 9     /* growth step (1-based) of stages 2 and 3 */
#define RHG_CREC_W1 10      /* first-bead Rosenbluth weight */
#define RHG_CREC_N 11

/* WidomAdsorbateMove of a flexible chain (src/mc_moves.c:7844-7946;
 * GrowMolecule src/cbmc.c:3056-3185 with HandleFirstBead :623-795, Rosen
 * :2762-2849, GenerateTrialOrientationsSimpleSphere :1480-1578,
 * GenerateTrialOrientationsMCScheme :1580-2482, ComputeExternalEnergies
 * :1154-1478) for insertions [first, first+n) of a job of n_total, as
 * rhg_widom_run does for rigid molecules: LJ biasing, framework VDW (grids
 * in use are read), infinite dilution.  Every insertion grows its internal
 * Metropolis walk from the 120 degree cone with 4 x moves_per_open_bead
 * moves per branch (the reference's start without a stored configuration,
 * src/cbmc.c:1662-1666, 1797-1800).  Draws come from the insertion's
 * counter stream, one sub-stream per step and purpose.  Blocks are ADDED
 * to *blocks; rec (nullable) receives n x RHG_CREC_N doubles, xyz
 * (nullable) the grown molecules, n x n_atoms x 3 (zero when rejected). */
int rhg_widom_chain_run(rhg_ctx *ctx, const rhg_chain *chain, uint64_t seed, int64_t n_total, int64_t first,
                        int64_t n, int32_t steps_per_cycle, rhg_widom_blocks *blocks, double *rec,
                        double *xyz);

/* ----------------------------------------------------------------- grids */
/* MakeGrid (src/grids.c:431-583) for one probe type over a unit-cell box.
 * fp64 = 0 stores RASPA2's FP32 node layout [i][j][k][8]; fp64 = 1 keeps the
 * same 8 channels in FP64. */
int rhg_make_grid(rhg_ctx *ctx, int32_t type, const double unit_box[9], double spacing,
                  int32_t fp64, rhg_grid **out);
/* n[3], size[3], shift[3], delta[3] as doubles, then bytes per node */
int rhg_grid_info(const rhg_grid *g, double out[13]);
int rhg_grid_download(const rhg_grid *g, void *host, size_t bytes);
/* REAL InterpolateVDWGrid(int,VECTOR) (src/grids.h:70, src/grids.c:979) */
int rhg_interpolate_vdw_grid(rhg_grid *g, int64_t n, const double *xyz, double *u);
/* Grids run on their context's stream: use them only while it lives.
 * rhg_grid_free itself needs no live context. */
void rhg_grid_free(rhg_grid *g);
/* void CalculateDerivativesAtPositionVDW(VECTOR,int,REAL*,VECTOR*,REAL_MATRIX3x3*,REAL*)
 * (src/framework_hessian.h:43, src/framework_hessian.c:63): out n x 14
 * (U, dU[3], H[9] in ax..cz order, third). */
int rhg_derivatives_at_positions_vdw(rhg_ctx *ctx, int32_t type, int64_t n, const double *xyz,
                                     double *out);
/* MakeGrid's Ewald-Coulomb part (src/grids.c:587-689): the real-space Ewald
 * potential of a unit charge (pseudo-atom "UNIT"), capped where it exceeds
 * the overlap criterion or an atom is closer than 1 A.  Needs the Ewald
 * charge method (other methods: RHG_ERR_UNSUPPORTED). */
int rhg_make_coulomb_grid(rhg_ctx *ctx, const double unit_box[9], double spacing, int32_t fp64,
                          rhg_grid **out);
/* REAL InterpolateCoulombGrid(int,VECTOR) (src/grids.c:1200-1295): the
 * tricubic value (no overlap short-cut) times the probe charge */
int rhg_interpolate_coulomb_grid(rhg_grid *g, int64_t n, const double *xyz, double charge, double *u);
/* REAL CalculateDerivativesAtPositionReal(VECTOR,int,REAL*,VECTOR*,REAL_MATRIX3x3*,REAL*)
 * (src/framework_hessian.c:111-175) for a unit charge: out n x 14 as the
 * VDW form, smallest_r[n] its return value (DBL_MAX when no atom is in range). */
int rhg_derivatives_at_positions_real(rhg_ctx *ctx, int64_t n, const double *xyz, double *out,
                                      double *smallest_r);
/* WriteVDWGrid / WriteCoulombGrid (src/grids.c:692-742, 815-876): RASPA2's
 * binary .grid file of an FP32 grid.  n_unit_cells null writes the number of
 * unit cells of the context's box.  Errors: RHG_ERR_IO. */
int rhg_grid_write(const rhg_grid *g, const char *path, const int32_t n_unit_cells[3]);
/* ReadVDWGrid / ReadCoulombGrid (src/grids.c:745-813, 879-975): a .grid file
 * onto the device of ctx.  The unit cell's vector lengths must match the
 * file's within 1e-8 A and, for a Coulomb grid with n_unit_cells given, the
 * number of unit cells exactly (RHG_ERR_INVALID).  A file whose length does
 * not match its header is RHG_ERR_IO. */
int rhg_grid_read(rhg_ctx *ctx, const char *path, int32_t coulomb, const double unit_box[9],
                  const int32_t n_unit_cells[3], rhg_grid **out);
/* spacing, unit-cell lengths[3], number of unit cells[3], coulomb flag,
 * Ewald precision of a Coulomb grid (-1 otherwise), (node, atom) pair
 * evaluations the build executed (through the cell lists; all pairs when
 * built without them; 0 for a grid read from a file) */
int rhg_grid_meta(const rhg_grid *g, double out[10]);
/* UseTabularGrid (src/input.c:5626-5630, 6190-6195): energies of pseudo-atom
 * `type` at a position, in rhg_framework_energy_at_positions and in the
 * Widom pipeline, interpolate its VDW grid (CalculateFrameworkVDWEnergyAtPosition's
 * grid branch, src/framework_energy.c:73-76); with a Coulomb grid in use the
 * real-space term of every charged pseudo-atom is the grid value times its
 * charge (src/framework_energy.c:197-200).  The grid is borrowed: it must
 * live, on the context's device, until it is replaced or detached (null). */
int rhg_use_grid(rhg_ctx *ctx, int32_t type, rhg_grid *g);
int rhg_use_coulomb_grid(rhg_ctx *ctx, rhg_grid *g);

/* ------------------------------------------------------------- profiling */
/* HIP-event timing of the kernels launched on the context's stream.
 * kernel families: "site_energy", "bin", "fourier", "widom_gen",
 * "widom_select", "block_reduce", "grid_build", "grid_interp", "sf", ...
 * The site-energy launches also count per site list, under
 * "site_energy:bfb" (Widom first-bead trials), "site_energy:bor"
 * (orientation trials of the other beads), "site_energy:bcc" (charged beads
 * of the chosen orientations), "site_energy:e" (energies at given
 * positions), "site_energy:cfb" and "site_energy:cgr" (chain first beads
 * and grown beads).
 * Returns launch count and summed milliseconds.  on: 0 off, 1 every family,
 * RHG_PROFILE_SITE_ENERGY only the site-energy kernels (each timed pair of
 * events adds a gap of some microseconds between launches, so timing only
 * the family of interest leaves the others back to back). */
#define RHG_PROFILE_SITE_ENERGY 2
int rhg_profile_enable(rhg_ctx *ctx, int32_t on);
int rhg_profile_get(rhg_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms);
int rhg_profile_reset(rhg_ctx *ctx);
int rhg_synchronize(rhg_ctx *ctx);
/* No counterpart in the reference: `reps` launches of a plain streaming
 * copy kernel (k_stream_copy) over `bytes` (multiple of 16), which reads and
 * writes exactly `bytes` each; a profiler's byte counters on it give the
 * scale to turn counter readings into bytes. */
int rhg_stream_copy(int device, int64_t bytes, int32_t reps);
/* Work done by rhg_widom_run since the last reset (no counterpart in the
 * reference; feeds the roofline): out[0] insertions, [1] first-bead sites,
 * [2] orientation sites, [3] Coulomb sites of selected orientations, [4]
 * Fourier molecules, [5] LJ pair evaluations, [6] LJ+Coulomb pair
 * evaluations, [7] Coulomb-only pair evaluations (all-pairs counts), [8]
 * k-vector x charged site terms, [9..11] pair evaluations executed by the
 * culled kernels (every site-energy call): LJ, LJ+Coulomb, Coulomb. */
int rhg_widom_stats(rhg_ctx *ctx, double out[12], int32_t reset);
/* Cell-list culling of the site-energy kernels (analogue of RASPA2's
 * UseCellLists, src/framework_energy.c:78-126, which the reference leaves
 * disabled): sites are binned into cells of about cell_size A and each
 * wave walks the ascending list of atoms within reach of its cell.  The FP64
 * result is bit-identical to the all-pairs loop.  on = 0: all pairs; 1: cell
 * lists; 2 (default): cell lists, and in mixed precision per-cell lists of
 * atom images (FP32 offsets from the cell centre), so that no pair needs a
 * minimum-image step (used when every perpendicular width is >= 2 cut-offs).
 * Default cell 2.5 A. */
int rhg_set_culling(rhg_ctx *ctx, int32_t on, double cell_size);
/* Launch shape of the culled site kernels (no counterpart in the reference;
 * a tuning knob): sites per lane of the LJ-only image-list kernels (1..4:
 * tiles of 64 x that many sites of one cell; default 4; lists with Coulomb
 * terms always take one) and waves per workgroup (1, 2 or 4; default 2).
 * Results do not depend on either. */
int rhg_set_tiles(rhg_ctx *ctx, int32_t sites_per_lane, int32_t waves_per_block);
/* Scratch buffers that contexts can borrow (no counterpart in the
 * reference): while a workspace is attached, the context's per-call device
 * buffers (Widom chunk buffers, binning scratch, Fourier partials) live in
 * it, so that many resident contexts served one at a time per stream share
 * one set of buffers.  Attaching a workspace owned by another context waits
 * for that context's stream and detaches it there; ws = NULL detaches. */
int rhg_workspace_create(int device, rhg_workspace **out);
void rhg_workspace_destroy(rhg_workspace *ws);
int rhg_workspace_footprint(const rhg_workspace *ws, size_t *bytes);
int rhg_attach_workspace(rhg_ctx *ctx, rhg_workspace *ws);

/* ---------------------------------------------------- screening service */
/* No counterpart in the reference, which answers one request per process
 * (python/raspa2.py:97-110; src/monte_carlo.c:626-748 runs the systems of
 * a process one after another).  A server holds a catalogue of frameworks,
 * keeps the ones in use resident on the device (atom records, k-tables,
 * cell lists and, for grid-served frameworks, energy grids) within an HBM
 * budget, evicting idle ones least-recently-used first, and runs up to
 * `slots` Widom requests at once, each on its slot's stream, starting the
 * next arrived request as soon as a slot and the memory free up (continuous
 * batching).  Requests can be submitted before the request loop starts or
 * while it runs.  A request is enqueued whole, in chunks of at most `slice`
 * insertions whose block sums accumulate on the device (chunks after the
 * first replay one captured HIP graph), and waited for once; its result is
 * bitwise the same alone or beside others, and a request that fails is
 * discarded and run again (at most 3 times). */
typedef struct rhg_server rhg_server;

typedef struct {
  int32_t n_sites;
  double eps_k[RHG_MAX_SITES], sigma[RHG_MAX_SITES], charge[RHG_MAX_SITES];
  double body[RHG_MAX_SITES][3];
  int32_t start, k_fb, k_or;
} rhg_probe;

typedef struct {
  int64_t framework;     /* id given to rhg_server_add_framework */
  rhg_probe probe;
  uint64_t seed;
  int64_t n_insertions;
  int32_t steps_per_cycle;
  double arrival_s;      /* arrival, seconds after the request loop starts; a
                            request submitted while it runs arrives no earlier
                            than its submission */
} rhg_request;

typedef struct {
  int32_t status;        /* RHG_OK or the error that ended the request */
  int32_t retries;       /* times the request was run again after a failure */
  rhg_widom_blocks blocks;
  double queued_s, started_s, finished_s, latency_s; /* latency = finished - arrival */
} rhg_result;

/* hbm_budget 0 = 80% of the free device memory; precision RHG_PREC_*;
 * slice = largest chunk (insertions per pipeline pass, <= 2^24). */
int rhg_server_create(int device, size_t hbm_budget, int32_t slots, int32_t precision, int64_t slice,
                      rhg_server **out);
/* Stops the request loop (finishing every submitted request) and frees all. */
void rhg_server_destroy(rhg_server *s);
const char *rhg_server_last_error(const rhg_server *s);
int rhg_server_add_framework(rhg_server *s, int64_t id, const rhg_framework *fw, const rhg_cell *cell,
                             const rhg_forcefield *ff);
/* Serves the requests of framework id from tabulated energy grids
 * (UseTabularGrid, src/input.c:5626-5630; src/framework_energy.c:73-76,
 * 197-200): a VDW grid per probe pseudo-atom and, for charged probes with
 * Ewald, the Ewald-Coulomb grid, built on first use over the framework's box
 * at `spacing` [A] (FP32 nodes, MakeGrid's layout), kept resident with the
 * framework, counted in admission and evicted with it; grids a request's
 * probe does not use are freed first when its footprint does not fit.
 * spacing = 0 returns to the direct sums. */
int rhg_server_use_grids(rhg_server *s, int64_t id, double spacing);
/* Per size bucket (frameworks of up to max_atoms[i] atoms): cell-list cell
 * size, chunk, and the site kernels' launch shape as rhg_set_tiles takes it:
 * sites per lane of the LJ-only lists (1..4) and waves per workgroup (1, 2,
 * 4) (nullable: 4 and 2), e.g. from an autotuner. */
int rhg_server_set_tuning(rhg_server *s, int32_t n, const int32_t *max_atoms, const double *cell_size,
                          const int64_t *chunk, const int32_t *sites_per_lane, const int32_t *waves_per_block);
/* Thread-safe; may be called while the request loop runs. */
int rhg_server_submit(rhg_server *s, const rhg_request *req, int64_t *id);
/* Host-side fault hook of a request not yet started (as rhg_inject_fault). */
int rhg_server_inject_fault(rhg_server *s, int64_t id, int64_t after_chunks);
/* Starts the request loop and returns; requests are served as they arrive. */
int rhg_server_start(rhg_server *s);
/* Waits until request id has finished: 1 finished, 0 not within timeout_s
 * (< 0: no limit), negative on error. */
int rhg_server_wait(rhg_server *s, int64_t id, double timeout_s);
/* Finishes every submitted request and stops the request loop. */
int rhg_server_stop(rhg_server *s);
/* rhg_server_start then rhg_server_stop: serves a trace of requests
 * submitted beforehand (arrivals are honoured) and those submitted meanwhile. */
int rhg_server_run(rhg_server *s);
/* RHG_OK once the request has finished (RHG_ERR_INVALID before). */
int rhg_server_result(rhg_server *s, int64_t id, rhg_result *out);
/* out[0] contexts created, [1] evictions, [2] peak resident bytes, [3]
 * resident bytes, [4] retries, [5] budget, [6] idle wait of the workers [s],
 * [7] requests, host seconds in [8] set-up, [9] submits, [10] collects,
 * [11] grids built, [12] grids evicted, [13] grid bytes resident, [14]
 * requests finished, [15] insertions run, [16..18] pair evaluations the
 * culled site kernels executed for them (LJ, LJ + Coulomb, Coulomb; as
 * rhg_widom_stats), [19] reserved */
int rhg_server_stats(rhg_server *s, double out[20]);
/* Device bytes a context of this framework holds after Widom insertions of
 * a probe of n_sites sites (n_charged charged, n_types distinct
 * pseudo-atoms) in chunks of `chunk` at the given precision (host-only
 * arithmetic; the admission estimate of the server). */
size_t rhg_estimate_footprint(const rhg_framework *fw, const rhg_cell *cell, const rhg_forcefield *ff,
                              int32_t n_types, int64_t chunk, int32_t n_sites, int32_t n_charged,
                              int32_t k_fb, int32_t k_or, double cell_size, int32_t precision);

/* ------------------------------------------------------ workload utility */
/* No counterpart in the reference (its frameworks come from CIF files):
 * synthetic framework positions for benchmark and test workloads, by random
 * sequential addition of uniform positions with a minimum-image minimum
 * separation d_min [A].  Returns RHG_ERR_INVALID if fewer than n atoms could
 * be placed within max_attempts (n_placed says how many). */
int rhg_synthetic_positions(uint64_t seed, const double box[9], int32_t n, double d_min,
                            int64_t max_attempts, double *xyz, int32_t *n_placed);

#ifdef __cplusplus
}
#endif
#endif
