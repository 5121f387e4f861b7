/* The RASPA2-side binding a maintainer adds to RASPA2 (context creation,
 * energies at a position, Widom moves), compiled against
 * include/rhg.h and linked with librhg.so.  RASPA2's globals are replaced by
 * a small host model with the same roles (framework atoms with pseudo-atom
 * types, the PotentialType / PotentialParms pair tables, the probe's
 * component definition, the Widom block accumulators), filled from a text
 * deck:
 *   n_atoms n_types boundary  box[9]  cutoff  temperature  charge_method
 *   n_types lines: eps_k sigma charge is_framework
 *   n_types x n_types lines: potential (1 = Lennard-Jones, 0 = none) eps_k sigma
 *   n_atoms lines: x y z type
 *   probe: n_sites start k_fb k_or, then n_sites lines: type x y z
 *   n_positions, then n_positions lines: x y z
 *
 *   binding <deck>              energies at the positions and a Widom run (GPU)
 *   binding <deck> --host-only  the host-side checks only (no GPU call)
 *
 * Energies are printed in the library's internal unit, one "u_vdw u_cc" line
 * per position, then the block sums "count sum_w sum_wdu ideal_gas" of each
 * block. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rhg.h"

#define MAX_T 16
#define LENNARD_JONES 1
#define NR_BLOCKS 5

static const double K_TO_E = 0.8314464919; /* KELVIN_TO_ENERGY, src/constants.c:180 */

typedef struct { double x, y, z; } POINT;
typedef struct { POINT Position; double Charge; int Type; } ATOM;
typedef struct { double Charge1; int FrameworkAtom; } PSEUDO_ATOM;

/* the host model (the roles of RASPA2's globals for one system) */
static int NumberOfAtoms, NumberOfPseudoAtoms, Boundary, ChargeEwald;
static ATOM Atoms[4096];
static PSEUDO_ATOM PseudoAtoms[MAX_T];
static int PotentialType[MAX_T][MAX_T];
static double PotentialParms[MAX_T][MAX_T][2]; /* eps (internal units), sigma */
static double Box[9], CutOff, Temperature;
static int ProbeN, ProbeStart, ProbeKfb, ProbeKor, ProbeType[RHG_MAX_SITES];
static POINT ProbePos[RHG_MAX_SITES];
static double WidomRosenbluthFactorCount[NR_BLOCKS], WidomRosenbluthFactorAccumulated[NR_BLOCKS],
    WidomEnergyDifferenceAccumulated[NR_BLOCKS], WidomIdealGasAccumulated[NR_BLOCKS];

static rhg_ctx *RhgContext;
static int32_t RhgType[MAX_T];

/* librhg mixes the framework atoms' and the probe's self parameters by
 * Lorentz-Berthelot; a force field whose pair table says otherwise (explicit
 * pair lines in force_field.def, src/potentials.c:3000-3130, or a potential
 * other than Lennard-Jones) is refused instead of silently changed.
 * Returns the number of pairs that differ. */
static int RhgCheckMixing(void)
{
  int bad = 0, a, t;
  for (a = 0; a < NumberOfPseudoAtoms; a++) {
    if (!PseudoAtoms[a].FrameworkAtom) continue;
    for (t = 0; t < NumberOfPseudoAtoms; t++) {
      if (PseudoAtoms[t].FrameworkAtom) continue;
      double e = sqrt(PotentialParms[a][a][0] * PotentialParms[t][t][0]);
      double s = 0.5 * (PotentialParms[a][a][1] + PotentialParms[t][t][1]);
      if (PotentialType[a][t] != LENNARD_JONES || PotentialType[a][a] != LENNARD_JONES ||
          PotentialType[t][t] != LENNARD_JONES || fabs(PotentialParms[a][t][0] - e) > 1e-12 * e ||
          fabs(PotentialParms[a][t][1] - s) > 1e-12 * s) {
        fprintf(stderr, "librhg: pair %d-%d is not a Lorentz-Berthelot mix of Lennard-Jones terms\n", a, t);
        bad++;
      }
    }
  }
  return bad;
}

/* context creation, once per system after ReadInput (src/input.c:8482-8488) */
static int RhgCreate(int device)
{
  int n = NumberOfAtoms, i, t;
  if (RhgCheckMixing()) return RHG_ERR_UNSUPPORTED;
  double *x = malloc(6 * (size_t)(n > 0 ? n : 1) * sizeof(double)), *y = x + n, *z = y + n, *q = z + n,
         *e = q + n, *sg = e + n;
  for (i = 0; i < n; i++) {
    ATOM *a = &Atoms[i];
    x[i] = a->Position.x; y[i] = a->Position.y; z[i] = a->Position.z; q[i] = a->Charge;
    e[i] = PotentialParms[a->Type][a->Type][0] / K_TO_E; /* back to K */
    sg[i] = PotentialParms[a->Type][a->Type][1];
  }
  rhg_framework fw = {n, x, y, z, q, e, sg, NULL};
  rhg_cell cell;
  memcpy(cell.box, Box, 9 * sizeof(double));
  cell.boundary = Boundary;
  rhg_forcefield ff = {CutOff, 1, ChargeEwald ? RHG_CHARGE_EWALD : RHG_CHARGE_NONE, CutOff, 1e-6, Temperature};
  int rc = rhg_create(&fw, &cell, &ff, device, &RhgContext);
  free(x);
  if (rc != RHG_OK) return rc;
  for (t = 0; t < NumberOfPseudoAtoms; t++)
    if (!PseudoAtoms[t].FrameworkAtom)
      if ((rc = rhg_add_pseudo_atom(RhgContext, PotentialParms[t][t][0] / K_TO_E, PotentialParms[t][t][1],
                                    PseudoAtoms[t].Charge1, &RhgType[t])) != RHG_OK)
        return rc;
  return rhg_set_precision(RhgContext, RHG_PREC_FP64);
}

/* CalculateFrameworkVDWEnergyAtPosition and
 * CalculateFrameworkChargeEnergyAtPosition (src/framework_energy.c:60, 176) */
static double CalculateFrameworkVDWEnergyAtPosition(POINT posA, int typeA, double scaling)
{
  double p[3] = {posA.x, posA.y, posA.z}, u = 0.0;
  rhg_framework_vdw_energy_at_position(RhgContext, p, RhgType[typeA], scaling, &u);
  return u;
}

static void CalculateFrameworkChargeEnergyAtPosition(POINT posA, int typeA, double *UChargeCharge,
                                                     double *UChargeBondDipole, double scaling)
{
  double p[3] = {posA.x, posA.y, posA.z};
  *UChargeCharge = *UChargeBondDipole = 0.0;
  if (ChargeEwald)
    rhg_framework_charge_energy_at_position(RhgContext, p, RhgType[typeA], scaling, UChargeCharge,
                                            UChargeBondDipole);
}

/* a Widom-only component's insertions (src/mc_moves.c:7844) */
static int RhgWidom(uint64_t seed, int64_t n_total, int steps_per_cycle)
{
  int i, k;
  rhg_molecule m;
  memset(&m, 0, sizeof(m));
  m.n_sites = ProbeN;
  for (i = 0; i < m.n_sites; i++) {
    m.type[i] = RhgType[ProbeType[i]];
    m.body[i][0] = ProbePos[i].x;
    m.body[i][1] = ProbePos[i].y;
    m.body[i][2] = ProbePos[i].z;
  }
  m.start = ProbeStart;
  m.k_fb = ProbeKfb;
  m.k_or = ProbeKor;
  rhg_widom_blocks b;
  memset(&b, 0, sizeof(b));
  int rc = rhg_widom_run(RhgContext, &m, seed, n_total, 0, n_total, steps_per_cycle, &b, NULL);
  if (rc != RHG_OK) return rc;
  for (k = 0; k < NR_BLOCKS; k++) {
    WidomRosenbluthFactorCount[k] += b.count[k];
    WidomRosenbluthFactorAccumulated[k] += b.sum_w[k];
    WidomEnergyDifferenceAccumulated[k] += b.sum_wdu[k];
    WidomIdealGasAccumulated[k] += b.ideal_gas[k];
  }
  return RHG_OK;
}

static int read_deck(const char *path, POINT **pos, int *npos)
{
  FILE *f = fopen(path, "r");
  int i, j, ok = 1;
  if (!f) return 0;
  ok &= fscanf(f, "%d %d %d", &NumberOfAtoms, &NumberOfPseudoAtoms, &Boundary) == 3;
  for (i = 0; i < 9; i++) ok &= fscanf(f, "%lf", &Box[i]) == 1;
  ok &= fscanf(f, "%lf %lf %d", &CutOff, &Temperature, &ChargeEwald) == 3;
  if (!ok || NumberOfAtoms > 4096 || NumberOfPseudoAtoms > MAX_T) { fclose(f); return 0; }
  for (i = 0; i < NumberOfPseudoAtoms; i++) {
    double e, s;
    ok &= fscanf(f, "%lf %lf %lf %d", &e, &s, &PseudoAtoms[i].Charge1, &PseudoAtoms[i].FrameworkAtom) == 4;
    PotentialParms[i][i][0] = e * K_TO_E;
    PotentialParms[i][i][1] = s;
  }
  for (i = 0; i < NumberOfPseudoAtoms; i++)
    for (j = 0; j < NumberOfPseudoAtoms; j++) {
      double e, s;
      ok &= fscanf(f, "%d %lf %lf", &PotentialType[i][j], &e, &s) == 3;
      PotentialParms[i][j][0] = e * K_TO_E;
      PotentialParms[i][j][1] = s;
    }
  for (i = 0; i < NumberOfAtoms; i++) {
    ATOM *a = &Atoms[i];
    ok &= fscanf(f, "%lf %lf %lf %d", &a->Position.x, &a->Position.y, &a->Position.z, &a->Type) == 4;
    a->Charge = PseudoAtoms[a->Type].Charge1;
  }
  ok &= fscanf(f, "%d %d %d %d", &ProbeN, &ProbeStart, &ProbeKfb, &ProbeKor) == 4;
  for (i = 0; i < ProbeN && i < RHG_MAX_SITES; i++)
    ok &= fscanf(f, "%d %lf %lf %lf", &ProbeType[i], &ProbePos[i].x, &ProbePos[i].y, &ProbePos[i].z) == 4;
  ok &= fscanf(f, "%d", npos) == 1;
  *pos = malloc(sizeof(POINT) * (size_t)(*npos > 0 ? *npos : 1));
  for (i = 0; i < *npos; i++) ok &= fscanf(f, "%lf %lf %lf", &(*pos)[i].x, &(*pos)[i].y, &(*pos)[i].z) == 3;
  fclose(f);
  return ok;
}

int main(int argc, char **argv)
{
  POINT *pos = NULL;
  int npos = 0, i, k;
  if (argc < 2 || !read_deck(argv[1], &pos, &npos)) {
    fprintf(stderr, "usage: binding <deck> [--host-only]\n");
    return 2;
  }
  if (argc > 2 && strcmp(argv[2], "--host-only") == 0) {
    /* host-only: the library loads, the mixing check decides, estimators run */
    rhg_widom_blocks b;
    double out[16];
    memset(&b, 0, sizeof(b));
    for (k = 0; k < NR_BLOCKS; k++) b.count[k] = b.ideal_gas[k] = 100.0, b.sum_w[k] = 1.0 + k;
    printf("version %s\n", rhg_version());
    printf("mixing %s\n", RhgCheckMixing() ? "refused" : "ok");
    const int rc = rhg_widom_estimators(&b, Temperature, 1000.0, 1.0, out);
    printf("estimators %d %.17g\n", rc, out[15]);
    free(pos);
    return 0;
  }
  int rc = RhgCreate(0);
  if (rc != RHG_OK) {
    fprintf(stderr, "RhgCreate: %d %s\n", rc, RhgContext ? rhg_last_error(RhgContext) : "");
    free(pos);
    return rc == RHG_ERR_UNSUPPORTED ? 3 : 1;
  }
  int probe_type = ProbeType[ProbeStart];
  for (i = 0; i < npos; i++) {
    double ucc, ucbd;
    double u = CalculateFrameworkVDWEnergyAtPosition(pos[i], probe_type, 1.0);
    CalculateFrameworkChargeEnergyAtPosition(pos[i], probe_type, &ucc, &ucbd, 1.0);
    printf("%.17g %.17g\n", u, ucc);
  }
  if ((rc = RhgWidom(4242, 20000, 20)) != RHG_OK) {
    fprintf(stderr, "RhgWidom: %d %s\n", rc, rhg_last_error(RhgContext));
    return 1;
  }
  for (k = 0; k < NR_BLOCKS; k++)
    printf("%.17g %.17g %.17g %.17g\n", WidomRosenbluthFactorCount[k], WidomRosenbluthFactorAccumulated[k],
           WidomEnergyDifferenceAccumulated[k], WidomIdealGasAccumulated[k]);
  rhg_destroy(RhgContext);
  free(pos);
  return 0;
}
