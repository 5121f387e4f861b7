/*
 * rhg_oracle.h -- CPU FP64 restatement of RASPA2's host-guest energy path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, and by nothing else.  The product path (raspa2_amd/)
 * never calls it.
 *
 * Every function restates the reference algorithm in plain C with strict
 * IEEE double arithmetic (built with -ffp-contract=off, no fast-math) and the
 * reference's operation order, so that values are bit-identical to what the
 * reference computes on the same inputs wherever the reference's arithmetic
 * is itself deterministic (libm calls aside).  Each function cites the
 * reference file:line it follows.
 *
 * Pins (tests/test_oracle_golden.py, tests/test_oracle_math.py): units,
 * mixing and shifts, cell volume and density, and the Henry estimator against
 * numbers printed in the reference's IRMOF-1 example output; MT19937-64
 * against std::mt19937_64; LJ and minimum image against closed forms and a
 * brute-force image sum; Ewald against the rock-salt Madelung constant;
 * Fourier Delta against the change of the full Fourier energy; tricubic
 * interpolation against exact tricubic polynomials; grid derivatives against
 * finite differences; Widom against the reference's He/IRMOF-1 result;
 * flexible-chain growth against hand-written growth plans, closed-form
 * bonded energies and empty-box quadrature of the torsion weight, bend-angle
 * and bond-length distributions (tests/test_oracle_chain.py).
 * Parity unpinned: the libm-dependent last bits (erfc, erf, exp, sin, cos)
 * and the global interleaving of a whole RASPA2 run's MT19937-64 stream
 * (orc_widom_run_mt's absolute offset).
 */
#ifndef RHG_ORACLE_H
#define RHG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_BC_RECTANGULAR 1
#define ORC_BC_TRICLINIC   2

#define ORC_CHARGE_NONE  0
#define ORC_CHARGE_EWALD 1

#define ORC_MAX_TYPES 64
#define ORC_MAX_SITES 8

/* ---------------------------------------------------------------- units */
/* out[0] energy unit [J], out[1] Coulomb conversion [internal*A/e^2],
 * out[2] energy->K, out[3] K->energy, out[4] k_B [internal/K],
 * out[5] Coulomb potential unit [K]  (src/constants.c:141-180) */
void orc_units(double out[6]);

/* ------------------------------------------------------------------ RNG */
/* MT19937-64 (src/utils.c:308-392) */
typedef struct { uint64_t mt[312]; int mti; } orc_mt64;
void     orc_mt64_seed(orc_mt64 *g, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *g);
double   orc_mt64_uniform(orc_mt64 *g);             /* RandomNumber(): closed [0,1] */
void     orc_mt64_sequence(uint64_t seed, int64_t n, uint64_t *out);
void     orc_mt64_uniform_sequence(uint64_t seed, int64_t n, double *out);

/* Counter-based per-insertion stream (SplitMix64 output function applied
 * to key + (k+1)*golden), mapped to [0,1] with RandomNumber()'s formula. */
uint64_t orc_stream_key(uint64_t seed, uint64_t insertion);
double   orc_stream_uniform(uint64_t key, uint64_t counter);
void     orc_stream_uniforms(uint64_t seed, uint64_t insertion, int64_t n, double *out);
void     orc_sincos_portable(double x, double *s, double *c);

/* ----------------------------------------------------------------- cell */
/* box/inv use RASPA's REAL_MATRIX3x3 field order: ax ay az bx by bz cx cy cz
 * (columns are the cell vectors a, b, c).  props: lengths[3], perpendicular
 * widths[3], volume. */
void orc_cell_make(double a, double b, double c, double alpha_deg, double beta_deg,
                   double gamma_deg, double box[9], double inv[9], double props[7]);
void orc_invert3x3(const double in[9], double out[9], double *det);
void orc_cell_properties(const double box[9], double props[7]);
void orc_min_image(const double box[9], const double inv[9], int bc, double dr[3]);

/* --------------------------------------------------------------- system */
typedef struct orc_system orc_system;

orc_system *orc_system_create(const double box[9], int bc, int n_atoms,
                              const double *x, const double *y, const double *z,
                              const double *charge, const double *eps_k, const double *sigma,
                              const double *mass,
                              double cutoff_vdw, int shifted,
                              int charge_method, double cutoff_charge, double ewald_precision,
                              double temperature);
void orc_system_destroy(orc_system *s);
/* registers a probe pseudo-atom; returns its type id */
int  orc_add_pseudo_atom(orc_system *s, double eps_k, double sigma, double charge);
/* blocked pockets: n centres (xyz, n x 3) with radii, periodic with the unit
 * cell unit_box; invert = 1 blocks everything outside the pockets
 * (src/framework.c:6168-6236) */
int orc_set_block_pockets(orc_system *s, int n, const double *xyz, const double *radius,
                          const double unit_box[9], int invert);
/* int BlockedPocket(VECTOR) (src/framework.c:6246-6300) */
int orc_blocked_pocket(const orc_system *s, const double pos[3]);
/* pair parameters of (type, atom): eps [internal], sigma [A], shift [internal] */
void orc_pair_params(const orc_system *s, int type, int atom, double out[3]);
void orc_system_info(const orc_system *s, double out[8]);  /* alpha,kx,ky,kz,nk,uion,beta,volume */

/* CalculateFrameworkVDWEnergyAtPosition, plain all-pairs branch */
double orc_vdw_energy_at(const orc_system *s, int type, const double pos[3], double scaling);
/* CalculateFrameworkChargeEnergyAtPosition, plain branch (returns Ucc) */
double orc_charge_energy_at(const orc_system *s, int type, const double pos[3], double scaling);
void   orc_energy_at_many(const orc_system *s, int type, int64_t n, const double *xyz,
                          double *u_vdw, double *u_cc);

/* ---------------------------------------------------------------- Ewald */
int    orc_ewald_nk(const orc_system *s);
void   orc_ewald_tables(const orc_system *s, double *kvec3, double *kfactor,
                        double *sf_re, double *sf_im, int *kidx3);
double orc_ewald_fourier_self_energy_framework(const orc_system *s); /* sum kf |S_f|^2 */
double orc_ewald_real_energy_framework(const orc_system *s);         /* sum_{i<j} erfc */
double orc_ewald_self_energy_framework(const orc_system *s);
/* CalculateEwaldFourierAdsorbate(TRUE,FALSE,...) at zero loading for one molecule:
 * out[0] UHostAdsorbateChargeChargeFourierDelta, out[1] UAdsorbateAdsorbateChargeChargeFourierDelta,
 * out[2] energy_self_new, out[3] energy_excluded_new, out[4] sum_k kf |S_new|^2 */
void   orc_fourier_adsorbate(const orc_system *s, int n_sites, const int *types,
                             const double *xyz, double out[5]);

/* ---------------------------------------------------------------- Widom */
typedef struct {
  int n_sites;
  int type[ORC_MAX_SITES];
  double body[ORC_MAX_SITES][3];
  int start;          /* StartingBead */
  int k_fb, k_or;     /* NumberOfTrialPositionsForTheFirstBeadWidom / NumberOfTrialPositionsWidom */
} orc_molecule;

/* per-insertion result record (doubles), see ORC_REC_* */
#define ORC_REC_W        0   /* RosenbluthNew incl. Fourier factor (0 on overlap) */
#define ORC_REC_DU       1   /* DeltaU of WidomAdsorbateMove (0 on overlap) */
#define ORC_REC_UVDW     2
#define ORC_REC_UREAL    3
#define ORC_REC_UFHOST   4
#define ORC_REC_UFADS    5
#define ORC_REC_OVERLAP  6   /* 1 if the insertion was rejected (OVERLAP) */
#define ORC_REC_SELFB    7
#define ORC_REC_SELOR    8
#define ORC_REC_FBMASK   9   /* bitmask of overlapping first-bead trials */
#define ORC_REC_ORMASK  10   /* bitmask of overlapping orientation trials */
#define ORC_REC_W1      11
#define ORC_REC_W2      12
#define ORC_REC_STAGE   13   /* 0: fb all overlap, 1: fb W<min, 2: or all overlap, 3: or W<min, 4: ok, 5: blocked */
#define ORC_REC_UADSVDW 14   /* UAdsorbateVDWNew (guest-guest, finite loading) */
#define ORC_REC_UADSREAL 15  /* UAdsorbateChargeChargeNew (real space) */
#define ORC_REC_N       16

/* trial record: what the insertion drew, usable as host-supplied trials */
typedef struct {
  double fb[64][3];
  double cord[64][ORC_MAX_SITES][3];
  double u_sel_fb, u_sel_or;
} orc_trials;

typedef double (*orc_uniform_fn)(void *state);

/* One WidomAdsorbateMove (src/mc_moves.c:7844) using uniforms from `next`. */
void orc_widom_insertion(const orc_system *s, const orc_molecule *m, orc_uniform_fn next,
                         void *state, double rec[ORC_REC_N], orc_trials *trials_out);
/* Same, with all draws taken from explicitly supplied trials. */
void orc_widom_insertion_trials(const orc_system *s, const orc_molecule *m,
                                const orc_trials *t, double rec[ORC_REC_N]);

/* RetraceMolecule(CBMC_DELETION) of an existing molecule (src/cbmc.c:2928-3054):
 * old = its n_sites x 3 positions, t the other trials (entries 0 ignored),
 * exclude = its adsorbate index (-1: none) */
void orc_widom_retrace_trials(const orc_system *s, const orc_molecule *m, int exclude, const double *old,
                              const orc_trials *t, double rec[ORC_REC_N]);
/* Counter-stream Widom over insertions [first, first+n): rec is n x ORC_REC_N */
void orc_widom_run_stream(const orc_system *s, const orc_molecule *m, uint64_t seed,
                          int64_t first, int64_t n, double *rec);
/* Trials of one insertion of the counter stream (for host-supplied GPU runs) */
void orc_stream_trials(const orc_system *s, const orc_molecule *m, uint64_t seed,
                       int64_t insertion, double *fb /*k_fb*3*/, double *cord /*k_or*n_sites*3*/,
                       double *u_sel /*2*/);

/* RASPA's own MC driver for a Widom-only single-component run, with the
 * MT19937-64 stream in RASPA's global draw order (src/monte_carlo.c:569-760).
 * blocks: 5 x 4 (count, sumW, sumW*dU, idealgas) ; returns #insertions */
int64_t orc_widom_run_mt(const orc_system *s, const orc_molecule *m, uint64_t seed,
                         int64_t n_cycles, int steps_per_cycle, double *blocks);
/* The same run, also returning for its first max_ins insertions the record
 * and the trials each drew (first-bead positions k_fb x 3, rotated offsets
 * k_or x n_sites x 3 from the starting bead, the two selection uniforms; the
 * layout of host-supplied trials; zero where the move drew nothing). Any
 * output may be NULL. */
int64_t orc_widom_run_mt_trials(const orc_system *s, const orc_molecule *m, uint64_t seed, int64_t n_cycles,
                                int steps_per_cycle, double *blocks, int64_t max_ins, double *recs,
                                double *fb, double *cord, double *u_sel);

/* Block index of each cycle (src/statistics.c:904,1126) */
void orc_block_of_cycles(int64_t n_cycles, int32_t *block);

/* estimators from 5-block accumulators (count,sumW,sumWdU,idealgas):
 * out[0..4] K_H per block, out[5] avg, out[6] 2sigma,
 * out[7..11] <U_gh>_1-<U_h>_0 per block [K], out[12] avg, out[13] 2sigma,
 * out[14] GetAverageHenryCoefficient, out[15] <W> overall */
void orc_widom_estimators(const double *blocks, double temperature, double density_kg_m3,
                          double ideal_gas_weight, double out[16]);
double orc_framework_density(double mass_g_mol, double volume_A3);

/* ----------------------------------------------------------------- grid */
/* MakeGrid geometry from a unit-cell box: out: N[3] (as doubles), size[3], shift[3], delta[3] */
void orc_grid_geometry(const double unit_box[9], double spacing, double out[12]);
/* CalculateDerivativesAtPositionVDW: out[0] U, out[1..3] dU, out[4..12] H (ax,ay,az,bx,by,bz,cx,cy,cz), out[13] third */
void orc_derivatives_vdw_at(const orc_system *s, int type, const double pos[3], double out[14]);
/* one stored node of MakeGrid (capped, scaled, FP32) */
void orc_grid_node(const orc_system *s, int type, const double geom[12], int i, int j, int k,
                   float out[8]);
/* CalculateDerivativesAtPositionReal (src/framework_hessian.c:111) for a
 * unit charge; returns the smallest in-range distance */
double orc_derivatives_real_at(const orc_system *s, const double pos[3], double out[14]);
/* UseTabularGrid: energies at a position of `type` (VDW) or of every
 * charged type (Coulomb) come from the grid (borrowed, [i][j][k][8] floats;
 * null detaches), src/framework_energy.c:73-76, 197-200 */
int orc_use_grid(orc_system *s, int type, const float *grid, const double geom[12], const double unit_box[9]);
int orc_use_coulomb_grid(orc_system *s, const float *grid, const double geom[12], const double unit_box[9]);
/* Finite loading: the adsorbate molecules present, molecule m holding
 * atoms [start[m], start[m+1]) of pseudo-atom types[] at xyz (n x 3).  The
 * adsorbate structure factor (StoreTotalChargeAdsorbates) and net charge are
 * computed from them (src/ewald.c:1289-1731).  n_mol = 0 clears. */
int orc_set_adsorbates(orc_system *s, int n_mol, const int *start, const int *types, const double *xyz);
/* CalculateInterVDWEnergyAdsorbateAtPosition (src/inter_energy.c:321) */
double orc_inter_vdw_at(const orc_system *s, int type, const double pos[3], int exclude, double scaling);
/* CalculateInterChargeEnergyAdsorbateAtPosition (src/inter_energy.c:390), chargeA = the
 * caller's scaling x charge */
double orc_inter_charge_at(const orc_system *s, int type, const double pos[3], int exclude, double charge_a);
/* CalculateEwaldFourierAdsorbate(New, Old, mol) (src/ewald.c:6859-7565): the
 * new molecule's n_new sites (0: none) and/or adsorbate molecule old_mol
 * (-1: none).  out: host delta, adsorbate delta, self(new-old),
 * excluded(new-old), sum_k kf(|S_ads'|^2-|S_ads|^2) */
void orc_fourier_move(const orc_system *s, int n_new, const int *new_types, const double *new_xyz,
                      int old_mol, double out[5]);
/* SurfaceAreaMove (src/mc_moves.c:16434-16520): area [A^2] and accessible
 * points per atom (nullable) with counter-stream points */
double orc_surface_area(const orc_system *s, int probe_type, double factor, int64_t samples, uint64_t seed,
                        const uint8_t *interacts /* nullable: all */, uint64_t *counts);
/* InterpolateCoulombGrid (src/grids.c:1200-1295) */
double orc_interpolate_coulomb_grid(const float *grid, const double geom[12], int bc, const double ub[9],
                                    const double uinv[9], const double pos[3], double charge);
/* Ewald-Coulomb grid node (src/grids.c:634-683) */
void orc_coulomb_grid_node(const orc_system *s, const double geom[12], int i, int j, int k, float out[8]);
/* Lekien-Marsden 64x64 matrix generated from its definition */
void orc_lm_coeff(int out[64 * 64]);
/* InterpolateVDWGrid with an explicit full grid in [i][j][k][8] float layout */
double orc_interpolate_grid(const float *grid, const double geom[12], int bc,
                            const double unit_box[9], const double unit_inv[9],
                            const double pos[3]);
/* The same interpolation with the 8 corner nodes computed on the fly
 * (orc_grid_node), for grids too large to hold on the host */
double orc_interpolate_direct(const orc_system *s, int type, const double geom[12], int bc,
                              const double unit_box[9], const double unit_inv[9],
                              const double pos[3]);

/* ------------------------------------------------------ flexible chains */
/* A flexible molecule of one flexible group (ReadComponentDefinition,
 * src/molecule.c:1430-3030): bonded terms with their parameters in the
 * units of the molecule file (K, A, degrees), converted as the reference
 * does on reading.  Kinds: bonds 0 FIXED_BOND (p0 = length), 1
 * HARMONIC_BOND; bends 0 HARMONIC_BEND; torsions 0 TRAPPE_DIHEDRAL. */
#define ORC_CHAIN_MAX 64
#define ORC_CHAIN_MAX_INTRA 2016
typedef struct {
  int n_atoms, start;
  int type[ORC_CHAIN_MAX];
  int n_bonds, bond[ORC_CHAIN_MAX][3];          /* a, b, kind */
  double bond_par[ORC_CHAIN_MAX][4];
  int n_bends, bend[2 * ORC_CHAIN_MAX][4];      /* a, b, c, kind */
  double bend_par[2 * ORC_CHAIN_MAX][4];
  int n_torsions, torsion[4 * ORC_CHAIN_MAX][5]; /* a, b, c, d, kind */
  double torsion_par[4 * ORC_CHAIN_MAX][4];
  int n_intra, intra[ORC_CHAIN_MAX_INTRA][2];
  double intra14_scaling;                       /* Intra14VDWScalingValue */
  int k_fb, k_trial, k_torsion, moves;          /* ...FirstBeadWidom, ...Widom, ...Torsion, ...PerOpenBead */
  double max_bond[ORC_CHAIN_MAX], max_bend[ORC_CHAIN_MAX], max_cone[ORC_CHAIN_MAX];
} orc_chain;

/* growth plan (SetConnectivityMatrix/SetGrowingStatus/Interactions,
 * src/cbmc.c:2582-2760, 901-1152): per step i (n_atoms - 1 of them)
 * info[8*i..] = CurrentBead, PreviousBead (-1), NumberOfPreviousBeads,
 * NumberOfBeadsToBePlaced, NumberOfBonds, NumberOfBends, NumberOfTorsions,
 * NumberOfIntraVDW; place[ORC_CHAIN_MAX*i..] the beads to be placed.
 * Returns the number of steps, or -1 for a molecule outside the supported
 * growth (a step with more than one previous bead). */
int orc_chain_plan(const orc_chain *c, int *info, int *place);

/* per-insertion chain record */
#define ORC_CREC_W        0  /* RosenbluthNew (0 on overlap or blocked) */
#define ORC_CREC_DU       1  /* DeltaU of WidomAdsorbateMove */
#define ORC_CREC_UHOSTVDW 2
#define ORC_CREC_UINTRA   3  /* UIntraVDWNew */
#define ORC_CREC_UBOND    4
#define ORC_CREC_UBEND    5
#define ORC_CREC_UTORSION 6
#define ORC_CREC_OVERLAP  7
#define ORC_CREC_STAGE    8  /* 0 fb overlap, 1 fb W<min, 2 step overlap, 3 step W<min, 4 ok, 5 blocked */
#define ORC_CREC_STEP     9  /* growth step of stages 2 and 3 (1-based) */
#define ORC_CREC_W1      10  /* first-bead weight */
#define ORC_CREC_N       11

/* WidomAdsorbateMove of a flexible chain (src/mc_moves.c:7844-7946,
 * GrowMolecule src/cbmc.c:3056-3185, Rosen :2762-2849, HandleFirstBead
 * :623-795, GenerateTrialOrientationsSimpleSphere :1480-1578,
 * GenerateTrialOrientationsMCScheme :1580-2482, ComputeExternalEnergies
 * :1154-1478) with the draws of insertion `insertion` of the counter stream
 * of `seed` (one sub-stream per purpose, see orc_chain_substream); xyz
 * (nullable) receives the grown molecule, n_atoms x 3. */
void orc_chain_insertion_stream(const orc_system *s, const orc_chain *c, uint64_t seed, int64_t insertion,
                                double rec[ORC_CREC_N], double *xyz);
/* insertions [first, first+n) of the counter stream: rec n x ORC_CREC_N,
 * xyz (nullable) n x n_atoms x 3 */
void orc_chain_run_stream(const orc_system *s, const orc_chain *c, uint64_t seed, int64_t first, int64_t n,
                          double *rec, double *xyz);
/* n insertions drawing from one MT19937-64 stream of `seed` in the
 * reference's order */
void orc_chain_run_mt(const orc_system *s, const orc_chain *c, uint64_t seed, int64_t n, double *rec,
                      double *xyz);
/* the same move drawing from one MT19937-64 stream in the reference's order */
void orc_chain_insertion_mt(const orc_system *s, const orc_chain *c, orc_mt64 *g, double rec[ORC_CREC_N],
                            double *xyz);
/* key of sub-stream `tag` of an insertion: tag = (step << 32) | j with j = 0
 * the internal MC scheme of the step, 1 its trial selection, 2 + t trial t */
uint64_t orc_chain_substream(uint64_t insertion_key, uint64_t tag);
/* bonded energies as the reference evaluates them [internal units]:
 * CalculateBondEnergy (src/internal_energy.c:372-490) of length r,
 * CalculateBendEnergy (:1168-1340) of atoms a-b-c, CalculateTorsionEnergy
 * (:2177-2430) of atoms a-b-c-d */
double orc_chain_bond_energy(const orc_system *s, int kind, const double par[4], double r);
double orc_chain_bend_energy(const orc_system *s, int kind, const double par[4], const double a[3],
                             const double b[3], const double c[3]);
double orc_chain_torsion_energy(const orc_system *s, int kind, const double par[4], const double a[3],
                                const double b[3], const double c[3], const double d[3]);

/* ------------------------------------------------- pore-size distribution */
/* SamplePoreSizeDistribution (src/sample.c:3444-3560) with CheckSphereOverlap
 * (:3376-3440): factor = PSDProbeDistance (1 Sigma, 2^(1/6) Minimum), n_b
 * sphere centres per cycle (20 x NumberOfCycles), counter-stream points.
 * orc_psd_largest: LargestRadius of one cycle (-10: none); orc_psd_histogram
 * ADDS the cycles' cumulative counts to hist[bins] */
double orc_psd_largest(const orc_system *s, double factor, int64_t n_b, uint64_t seed, int64_t i);
void orc_psd_histogram(const orc_system *s, double factor, int64_t n_b, uint64_t seed, int64_t first,
                       int64_t n, int bins, double range, double *hist);

#ifdef __cplusplus
}
#endif
#endif
