/*
 * rhg_oracle.c -- CPU FP64 restatement of RASPA2's host-guest energy path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Built with strict IEEE double
 * arithmetic (-ffp-contract=off, no fast-math) and the reference's operation
 * order.  Citations are to the RASPA2 sources (src/...).
 */
#include "rhg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* strict IEEE: the Makefile passes -ffp-contract=off */

/* ===================================================================== */
/* constants (src/constants.h:39-62)                                      */
/* ===================================================================== */
#define ANGSTROM_ 1e-10
#define PICO_SECOND_ 1e-12
#define ATOMIC_MASS_UNIT_ 1.6605402e-27
#define ELECTRONIC_CHARGE_UNIT_ 1.60217733e-19
#define MOLAR_GAS_CONSTANT_ 8.314464919
#define BOLTZMANN_CONSTANT_ 1.380650324e-23
#define AVOGADRO_CONSTANT_ 6.0221419947e23
#define ELECTRIC_CONSTANT_ 8.8541878176e-12
#define ENERGY_OVERLAP_CRITERIA 1.0e7          /* src/input.c:367 */
#define MINIMUM_ROSENBLUTH_FACTOR 1.0e-150     /* src/input.c:368 */
#define NR_BLOCKS 5                            /* src/statistics.h:40 */

#define SQR(x) ((x) * (x))
#define CUBE(x) ((x) * (x) * (x))

/* NINT of src/utils.h:60: round half away from zero through (int) */
static inline int nint(double x) { return (int)((x) >= 0.0 ? ((x) + 0.5) : ((x)-0.5)); }

/* src/constants.c:141-180 with LENGTH_UNIT=1e-10, TIME_UNIT=1e-12,
 * MASS_UNIT=amu, CHARGE_UNIT=e, DielectricConstantOfTheMedium=1 */
typedef struct { double E, C, E2K, K2E, KB; } units_t;
static units_t make_units(void) {
  volatile double length = ANGSTROM_, time = PICO_SECOND_, mass = ATOMIC_MASS_UNIT_,
                  charge = ELECTRONIC_CHARGE_UNIT_, dielectric = 1.0;
  units_t u;
  u.E = ((mass)*SQR(length) / SQR(time));
  u.KB = (BOLTZMANN_CONSTANT_ / u.E);
  u.C = (((SQR(charge) / (4.0 * M_PI * ELECTRIC_CONSTANT_ * length * u.E))) / (dielectric));
  u.E2K = ((u.E * AVOGADRO_CONSTANT_) / MOLAR_GAS_CONSTANT_);
  u.K2E = (MOLAR_GAS_CONSTANT_ / (u.E * AVOGADRO_CONSTANT_));
  return u;
}

void orc_units(double out[6]) {
  units_t u = make_units();
  out[0] = u.E;
  out[1] = u.C;
  out[2] = u.E2K;
  out[3] = u.K2E;
  out[4] = u.KB;
  out[5] = u.C * u.E2K;
}

/* ===================================================================== */
/* MT19937-64 (src/utils.c:308-392)                                      */
/* ===================================================================== */
#define MT_NN 312
#define MT_MM 156
#define MT_MATRIX_A 0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

void orc_mt64_seed(orc_mt64 *g, uint64_t seed) {
  g->mt[0] = seed;
  for (g->mti = 1; g->mti < MT_NN; g->mti++)
    g->mt[g->mti] = (6364136223846793005ULL * (g->mt[g->mti - 1] ^ (g->mt[g->mti - 1] >> 62)) +
                     (uint64_t)g->mti);
}

uint64_t orc_mt64_next(orc_mt64 *g) {
  static const uint64_t mag01[2] = {0ULL, MT_MATRIX_A};
  int i;
  uint64_t x;
  if (g->mti >= MT_NN) {
    if (g->mti == MT_NN + 1) orc_mt64_seed(g, 5489ULL);
    for (i = 0; i < MT_NN - MT_MM; i++) {
      x = (g->mt[i] & MT_UM) | (g->mt[i + 1] & MT_LM);
      g->mt[i] = g->mt[i + MT_MM] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    for (; i < MT_NN - 1; i++) {
      x = (g->mt[i] & MT_UM) | (g->mt[i + 1] & MT_LM);
      g->mt[i] = g->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    x = (g->mt[MT_NN - 1] & MT_UM) | (g->mt[0] & MT_LM);
    g->mt[MT_NN - 1] = g->mt[MT_MM - 1] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    g->mti = 0;
  }
  x = g->mt[g->mti++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= (x >> 43);
  return x;
}

/* RandomNumber(): closed interval [0,1] (src/utils.c:378-381) */
double orc_mt64_uniform(orc_mt64 *g) {
  return (double)(orc_mt64_next(g) >> 11) * (1.0 / 9007199254740991.0);
}

void orc_mt64_sequence(uint64_t seed, int64_t n, uint64_t *out) {
  orc_mt64 g;
  orc_mt64_seed(&g, seed);
  for (int64_t i = 0; i < n; i++) out[i] = orc_mt64_next(&g);
}

void orc_mt64_uniform_sequence(uint64_t seed, int64_t n, double *out) {
  orc_mt64 g;
  orc_mt64_seed(&g, seed);
  for (int64_t i = 0; i < n; i++) out[i] = orc_mt64_uniform(&g);
}

/* ===================================================================== */
/* Counter-based per-insertion stream (SplitMix64 output function,        */
/* Steele, Lea & Flood, OOPSLA 2014).  Not in the reference: it defines   */
/* the seeded trial positions of the data-parallel Widom path.            */
/* ===================================================================== */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
#define GOLDEN_GAMMA 0x9E3779B97F4A7C15ULL

/* sin and cos of x in [0, 2*pi] evaluated identically on host and device:
 * Cody-Waite reduction by pi/2 (two-part constant), Taylor series of sin
 * through r^21 and cos through r^20 on |r| <= pi/4, Horner form.  Used for
 * the Shoemake angles of counter-stream trials only. */
static void sincos_portable(double x, double *s, double *c) {
  const double two_over_pi = 0.63661977236758138243;
  const double pio2_hi = 1.57079632673412561417e+00;
  const double pio2_lo = 6.07710050650619224932e-11;
  double kd = trunc(x * two_over_pi + 0.5);
  int k = (int)kd;
  double r = (x - kd * pio2_hi) - kd * pio2_lo;
  double r2 = r * r;
  const double s3 = -1.0 / 6.0, s5 = 1.0 / 120.0, s7 = -1.0 / 5040.0, s9 = 1.0 / 362880.0,
               s11 = -1.0 / 39916800.0, s13 = 1.0 / 6227020800.0, s15 = -1.0 / 1307674368000.0,
               s17 = 1.0 / 355687428096000.0, s19 = -1.0 / 121645100408832000.0,
               s21 = 1.0 / 51090942171709440000.0;
  const double c2 = -0.5, c4 = 1.0 / 24.0, c6 = -1.0 / 720.0, c8 = 1.0 / 40320.0,
               c10 = -1.0 / 3628800.0, c12 = 1.0 / 479001600.0, c14 = -1.0 / 87178291200.0,
               c16 = 1.0 / 20922789888000.0, c18 = -1.0 / 6402373705728000.0,
               c20 = 1.0 / 2432902008176640000.0;
  double ps = s21;
  ps = ps * r2 + s19;
  ps = ps * r2 + s17;
  ps = ps * r2 + s15;
  ps = ps * r2 + s13;
  ps = ps * r2 + s11;
  ps = ps * r2 + s9;
  ps = ps * r2 + s7;
  ps = ps * r2 + s5;
  ps = ps * r2 + s3;
  double sr = r + (r * r2) * ps;
  double pc = c20;
  pc = pc * r2 + c18;
  pc = pc * r2 + c16;
  pc = pc * r2 + c14;
  pc = pc * r2 + c12;
  pc = pc * r2 + c10;
  pc = pc * r2 + c8;
  pc = pc * r2 + c6;
  pc = pc * r2 + c4;
  pc = pc * r2 + c2;
  double cr = 1.0 + r2 * pc;
  switch (k & 3) {
  case 0: *s = sr; *c = cr; break;
  case 1: *s = cr; *c = -sr; break;
  case 2: *s = -sr; *c = -cr; break;
  default: *s = -cr; *c = sr; break;
  }
}

void orc_sincos_portable(double x, double *s, double *c) { sincos_portable(x, s, c); }

uint64_t orc_stream_key(uint64_t seed, uint64_t insertion) {
  return mix64(seed ^ mix64(insertion + 0x632BE59BD9B4E019ULL));
}

double orc_stream_uniform(uint64_t key, uint64_t counter) {
  uint64_t x = mix64(key + (counter + 1ULL) * GOLDEN_GAMMA);
  return (double)(x >> 11) * (1.0 / 9007199254740991.0);
}

void orc_stream_uniforms(uint64_t seed, uint64_t insertion, int64_t n, double *out) {
  uint64_t key = orc_stream_key(seed, insertion);
  for (int64_t i = 0; i < n; i++) out[i] = orc_stream_uniform(key, (uint64_t)i);
}

/* ===================================================================== */
/* cell (src/framework.c:3024-3066, src/matrix.c:902-927,                 */
/*       src/framework.c:6338-6393)                                       */
/* ===================================================================== */
enum { AX = 0, AY, AZ, BX, BY, BZ, CX, CY, CZ };

void orc_invert3x3(const double in[9], double out[9], double *determinant) {
  double r, d;
  out[AX] = in[BY] * in[CZ] - in[BZ] * in[CY];
  out[AY] = in[AZ] * in[CY] - in[AY] * in[CZ];
  out[AZ] = in[AY] * in[BZ] - in[AZ] * in[BY];
  out[BX] = in[BZ] * in[CX] - in[BX] * in[CZ];
  out[BY] = in[AX] * in[CZ] - in[AZ] * in[CX];
  out[BZ] = in[AZ] * in[BX] - in[AX] * in[BZ];
  out[CX] = in[BX] * in[CY] - in[BY] * in[CX];
  out[CY] = in[AY] * in[CX] - in[AX] * in[CY];
  out[CZ] = in[AX] * in[BY] - in[AY] * in[BX];
  d = in[AX] * out[AX] + in[BX] * out[AY] + in[CX] * out[AZ];
  r = 0.0;
  if (fabs(d) > 0.0) r = 1.0 / d;
  out[AX] *= r; out[BX] *= r; out[CX] *= r;
  out[AY] *= r; out[BY] *= r; out[CY] *= r;
  out[AZ] *= r; out[BZ] *= r; out[CZ] *= r;
  *determinant = d;
}

void orc_cell_properties(const double in[9], double p[7]) {
  double axb1, axb2, axb3, bxc1, bxc2, bxc3, cxa1, cxa2, cxa3, vol;
  p[0] = sqrt(SQR(in[AX]) + SQR(in[AY]) + SQR(in[AZ]));
  p[1] = sqrt(SQR(in[BX]) + SQR(in[BY]) + SQR(in[BZ]));
  p[2] = sqrt(SQR(in[CX]) + SQR(in[CY]) + SQR(in[CZ]));
  axb1 = in[AY] * in[BZ] - in[AZ] * in[BY];
  axb2 = in[AZ] * in[BX] - in[AX] * in[BZ];
  axb3 = in[AX] * in[BY] - in[AY] * in[BX];
  bxc1 = in[BY] * in[CZ] - in[BZ] * in[CY];
  bxc2 = in[BZ] * in[CX] - in[BX] * in[CZ];
  bxc3 = in[BX] * in[CY] - in[BY] * in[CX];
  cxa1 = in[CY] * in[AZ] - in[AY] * in[CZ];
  cxa2 = in[AX] * in[CZ] - in[AZ] * in[CX];
  cxa3 = in[AY] * in[CX] - in[AX] * in[CY];
  vol = fabs(in[AX] * bxc1 + in[AY] * bxc2 + in[AZ] * bxc3);
  p[3] = vol / sqrt(bxc1 * bxc1 + bxc2 * bxc2 + bxc3 * bxc3);
  p[4] = vol / sqrt(cxa1 * cxa1 + cxa2 * cxa2 + cxa3 * cxa3);
  p[5] = vol / sqrt(axb1 * axb1 + axb2 * axb2 + axb3 * axb3);
  p[6] = vol;
}

void orc_cell_make(double A, double B, double C, double alpha_deg, double beta_deg,
                   double gamma_deg, double box[9], double inv[9], double props[7]) {
  double alpha = alpha_deg * (M_PI / 180.0);
  double beta = beta_deg * (M_PI / 180.0);
  double gamma = gamma_deg * (M_PI / 180.0);
  double det;
  double tempd = (cos(alpha) - cos(gamma) * cos(beta)) / sin(gamma);
  box[AX] = A;
  box[AY] = 0.0;
  box[AZ] = 0.0;
  box[BX] = B * cos(gamma);
  box[BY] = B * sin(gamma);
  box[BZ] = 0.0;
  box[CX] = C * cos(beta);
  box[CY] = C * tempd;
  box[CZ] = C * sqrt(1.0 - SQR(cos(beta)) - SQR(tempd));
  orc_invert3x3(box, inv, &det);
  orc_cell_properties(box, props);
}

/* ApplyBoundaryCondition (src/potentials.h:232-271) */
void orc_min_image(const double box[9], const double inv[9], int bc, double dr[3]) {
  double s[3], t[3];
  if (bc == ORC_BC_RECTANGULAR) {
    dr[0] -= box[AX] * (double)nint(dr[0] * inv[AX]);
    dr[1] -= box[BY] * (double)nint(dr[1] * inv[BY]);
    dr[2] -= box[CZ] * (double)nint(dr[2] * inv[CZ]);
  } else {
    s[0] = inv[AX] * dr[0] + inv[BX] * dr[1] + inv[CX] * dr[2];
    s[1] = inv[AY] * dr[0] + inv[BY] * dr[1] + inv[CY] * dr[2];
    s[2] = inv[AZ] * dr[0] + inv[BZ] * dr[1] + inv[CZ] * dr[2];
    t[0] = s[0] - (double)nint(s[0]);
    t[1] = s[1] - (double)nint(s[1]);
    t[2] = s[2] - (double)nint(s[2]);
    dr[0] = box[AX] * t[0] + box[BX] * t[1] + box[CX] * t[2];
    dr[1] = box[AY] * t[0] + box[BY] * t[1] + box[CY] * t[2];
    dr[2] = box[AZ] * t[0] + box[BZ] * t[1] + box[CZ] * t[2];
  }
}

/* ===================================================================== */
/* system                                                                */
/* ===================================================================== */
typedef struct { double re, im; } cplx;

struct orc_system {
  double box[9], inv[9], props[7];
  int bc;
  int n;
  double *x, *y, *z, *q, *eps, *sig;
  double mass_total;
  double rc_vdw2;
  int shifted;
  int charge_method;
  double rc_charge, rc_charge2;
  units_t u;
  double temperature, beta;
  int ntypes;
  double t_eps[ORC_MAX_TYPES], t_sig[ORC_MAX_TYPES], t_q[ORC_MAX_TYPES];
  double *pp[ORC_MAX_TYPES]; /* n x 3: eps_internal, sigma, shift */
  /* Ewald */
  double alpha;
  int kx, ky, kz, nk;
  double recip_cutoff;
  double *kvec, *kfactor;
  int *kidx;
  cplx *sf;
  double uion;
  double net_charge_framework;
  /* blocked pockets (src/framework.c:6168-6300): centres (xyz), radii, the
   * unit cell they repeat with, and the inverted mode */
  int n_pockets;
  double *pk; /* n x 4: x, y, z, radius */
  double pk_ubox[9], pk_uinv[9];
  int pk_invert;
  /* tabulated grids (FrameworkModel GRID, src/input.c:5626-5630): borrowed
   * node arrays [i][j][k][8] with their geometry, and the unit cell they
   * wrap positions into */
  const float *vgrid[ORC_MAX_TYPES];
  double vgeom[ORC_MAX_TYPES][12];
  const float *cgrid;
  double cgeom[12];
  double g_ubox[9], g_uinv[9];
  /* adsorbates present (finite loading) */
  int ads_nmol, ads_n;
  int *ads_start, *ads_type, *ads_mol;
  double *ads_xyz;
  cplx *sads;           /* StoreTotalChargeAdsorbates, k order */
  double net_charge_ads; /* NetChargeAdsorbates */
};

/* LENNARD_JONES branch of PotentialValue (src/potentials.c:5867-5877) */
static inline double lj_value(double arg1, double sigma, double arg3, double rr) {
  double arg2 = SQR(sigma);
  double rri3 = CUBE(arg2 / rr);
  return 4.0 * arg1 * (rri3 * (rri3 - 1.0)) - arg3;
}

int orc_set_block_pockets(orc_system *s, int n, const double *xyz, const double *radius,
                          const double unit_box[9], int invert) {
  free(s->pk);
  s->pk = NULL;
  s->n_pockets = 0;
  if (n <= 0) return 0;
  s->pk = (double *)malloc(sizeof(double) * 4 * (size_t)n);
  for (int i = 0; i < n; i++) {
    s->pk[4 * i] = xyz[3 * i];
    s->pk[4 * i + 1] = xyz[3 * i + 1];
    s->pk[4 * i + 2] = xyz[3 * i + 2];
    s->pk[4 * i + 3] = radius[i];
  }
  memcpy(s->pk_ubox, unit_box, 9 * sizeof(double));
  double det;
  orc_invert3x3(s->pk_ubox, s->pk_uinv, &det);
  s->pk_invert = invert ? 1 : 0;
  s->n_pockets = n;
  return 0;
}

/* ApplyBoundaryConditionUnitCell (src/potentials.c:5715-5752) */
static void min_image_unit_cell(const orc_system *s, double *dx, double *dy, double *dz) {
  const double *ub = s->pk_ubox, *ui = s->pk_uinv;
  if (s->bc == ORC_BC_RECTANGULAR) {
    /* UnitCellSize: the unit-cell edge lengths of a rectangular cell */
    *dx -= ub[AX] * (double)nint(*dx / ub[AX]);
    *dy -= ub[BY] * (double)nint(*dy / ub[BY]);
    *dz -= ub[CZ] * (double)nint(*dz / ub[CZ]);
  } else {
    double sx = ui[AX] * *dx + ui[BX] * *dy + ui[CX] * *dz;
    double sy = ui[AY] * *dx + ui[BY] * *dy + ui[CY] * *dz;
    double sz = ui[AZ] * *dx + ui[BZ] * *dy + ui[CZ] * *dz;
    double tx = sx - (double)nint(sx), ty = sy - (double)nint(sy), tz = sz - (double)nint(sz);
    *dx = ub[AX] * tx + ub[BX] * ty + ub[CX] * tz;
    *dy = ub[AY] * tx + ub[BY] * ty + ub[CY] * tz;
    *dz = ub[AZ] * tx + ub[BZ] * ty + ub[CZ] * tz;
  }
}

/* BlockedPocket (src/framework.c:6246-6300) */
int orc_blocked_pocket(const orc_system *s, const double pos[3]) {
  if (s->n_pockets <= 0) return 0;
  for (int i = 0; i < s->n_pockets; i++) {
    double dx = s->pk[4 * i] - pos[0], dy = s->pk[4 * i + 1] - pos[1], dz = s->pk[4 * i + 2] - pos[2];
    min_image_unit_cell(s, &dx, &dy, &dz);
    double r = sqrt(SQR(dx) + SQR(dy) + SQR(dz));
    if (r < s->pk[4 * i + 3]) return s->pk_invert ? 0 : 1;
  }
  return s->pk_invert ? 1 : 0;
}

int orc_add_pseudo_atom(orc_system *s, double eps_k, double sigma, double charge) {
  if (s->ntypes >= ORC_MAX_TYPES) return -1;
  int t = s->ntypes++;
  s->t_eps[t] = eps_k;
  s->t_sig[t] = sigma;
  s->t_q[t] = charge;
  s->pp[t] = (double *)malloc(sizeof(double) * 3 * (size_t)(s->n > 0 ? s->n : 1));
  double ei = eps_k * s->u.K2E; /* src/potentials.c:268-274 */
  for (int i = 0; i < s->n; i++) {
    double ej = s->eps[i] * s->u.K2E;
    double e = sqrt(ei * ej);                 /* src/potentials.c:3164-3171 */
    double sg = 0.5 * (sigma + s->sig[i]);
    double shift = 0.0;
    if (s->shifted) shift = lj_value(e, sg, 0.0, s->rc_vdw2); /* src/potentials.c:5263-5285 */
    s->pp[t][3 * i + 0] = e;
    s->pp[t][3 * i + 1] = sg;
    s->pp[t][3 * i + 2] = shift;
  }
  return t;
}

static void pair_lj(const orc_system *s, int ta, int tb, double out[3]) {
  double ei = s->t_eps[ta] * s->u.K2E, ej = s->t_eps[tb] * s->u.K2E;
  double e = sqrt(ei * ej);
  double sg = 0.5 * (s->t_sig[ta] + s->t_sig[tb]);
  out[0] = e;
  out[1] = sg;
  out[2] = s->shifted ? lj_value(e, sg, 0.0, s->rc_vdw2) : 0.0;
}

void orc_pair_params(const orc_system *s, int type, int atom, double out[3]) {
  out[0] = s->pp[type][3 * atom + 0];
  out[1] = s->pp[type][3 * atom + 1];
  out[2] = s->pp[type][3 * atom + 2];
}

static int considered_charged(double q) { return fabs(q) > 1e-10; }

/* InitializeEwald automatic branch (src/ewald.c:200-260) */
static void ewald_init(orc_system *s, double precision) {
  double eps = fmin(fabs(precision), 0.5);
  double rc = s->rc_charge;
  double tol = sqrt(fabs(log(eps * rc)));
  s->alpha = sqrt(fabs(log(eps * rc * tol))) / rc;
  double tol1 = sqrt(-log(eps * rc * SQR(2.0 * tol * s->alpha)));
  s->kx = (int)nint((double)0.25 + s->props[0] * s->alpha * tol1 / M_PI);
  s->ky = (int)nint((double)0.25 + s->props[1] * s->alpha * tol1 / M_PI);
  s->kz = (int)nint((double)0.25 + s->props[2] * s->alpha * tol1 / M_PI);
  int m = s->kx;
  if (s->ky > m) m = s->ky;
  if (s->kz > m) m = s->kz;
  s->recip_cutoff = SQR(1.05 * m);
  int nvec = 0;
  for (int ii = 0; ii <= s->kx; ii++)
    for (int jj = -s->ky; jj <= s->ky; jj++)
      for (int kk = -s->kz; kk <= s->kz; kk++) {
        double ksqr = SQR(ii) + SQR(jj) + SQR(kk);
        if ((ksqr != 0) && (ksqr < s->recip_cutoff)) nvec++;
      }
  s->nk = nvec;
}

/* SetupKVectors (src/ewald.c:532-600) */
static void ewald_setup_kvectors(orc_system *s) {
  s->kvec = (double *)malloc(sizeof(double) * 3 * (size_t)s->nk);
  s->kfactor = (double *)malloc(sizeof(double) * (size_t)s->nk);
  s->kidx = (int *)malloc(sizeof(int) * 3 * (size_t)s->nk);
  const double *inv = s->inv;
  double volume = s->props[6];
  double alpha = s->alpha;
  int nvec = 0;
  for (int ii = 0; ii <= s->kx; ii++) {
    double rk2x = 2.0 * M_PI * ii * inv[AX];
    double rk2y = 2.0 * M_PI * ii * inv[BX];
    double rk2z = 2.0 * M_PI * ii * inv[CX];
    for (int jj = -s->ky; jj <= s->ky; jj++) {
      double rk1x = rk2x + 2.0 * M_PI * jj * inv[AY];
      double rk1y = rk2y + 2.0 * M_PI * jj * inv[BY];
      double rk1z = rk2z + 2.0 * M_PI * jj * inv[CY];
      for (int kk = -s->kz; kk <= s->kz; kk++) {
        double rkx = rk1x + 2.0 * M_PI * kk * inv[AZ];
        double rky = rk1y + 2.0 * M_PI * kk * inv[BZ];
        double rkz = rk1z + 2.0 * M_PI * kk * inv[CZ];
        double ksqr = SQR(ii) + SQR(jj) + SQR(kk);
        if ((ksqr != 0) && (ksqr < s->recip_cutoff)) {
          s->kvec[3 * nvec + 0] = rkx;
          s->kvec[3 * nvec + 1] = rky;
          s->kvec[3 * nvec + 2] = rkz;
          s->kidx[3 * nvec + 0] = ii;
          s->kidx[3 * nvec + 1] = jj;
          s->kidx[3 * nvec + 2] = kk;
          double rksqr = SQR(rkx) + SQR(rky) + SQR(rkz);
          if (ii == 0)
            s->kfactor[nvec] = s->u.C * (2.0 * M_PI / volume) * exp((-0.25 / SQR(alpha)) * rksqr) / rksqr;
          else
            s->kfactor[nvec] = s->u.C * (4.0 * M_PI / volume) * exp((-0.25 / SQR(alpha)) * rksqr) / rksqr;
          nvec++;
        }
      }
    }
  }
}

/* exp(i k.r) tables for a set of sites, as in src/ewald.c:7045-7110.
 * eik{x,y,z}[j*ns + i] for j in [-kmax, kmax] stored with offset kmax. */
typedef struct { int ns, kx, ky, kz; cplx *ex, *ey, *ez; } eik_t;

static void eik_build(eik_t *e, int ns, const double (*fpos)[3], int kx, int ky, int kz) {
  e->ns = ns;
  e->kx = kx;
  e->ky = ky;
  e->kz = kz;
  e->ex = (cplx *)calloc((size_t)(2 * kx + 1) * (size_t)ns, sizeof(cplx));
  e->ey = (cplx *)calloc((size_t)(2 * ky + 1) * (size_t)ns, sizeof(cplx));
  e->ez = (cplx *)calloc((size_t)(2 * kz + 1) * (size_t)ns, sizeof(cplx));
#define EX(j, i) e->ex[((j) + kx) * ns + (i)]
#define EY(j, i) e->ey[((j) + ky) * ns + (i)]
#define EZ(j, i) e->ez[((j) + kz) * ns + (i)]
  for (int i = 0; i < ns; i++) {
    EX(0, i).re = 1.0; EX(0, i).im = 0.0;
    EY(0, i).re = 1.0; EY(0, i).im = 0.0;
    EZ(0, i).re = 1.0; EZ(0, i).im = 0.0;
    double px = fpos[i][0], py = fpos[i][1], pz = fpos[i][2];
    if (kx >= 1) { EX(1, i).re = cos(px); EX(1, i).im = sin(px); }
    if (ky >= 1) { EY(1, i).re = cos(py); EY(1, i).im = sin(py); EY(-1, i).re = cos(py); EY(-1, i).im = -sin(py); }
    if (kz >= 1) { EZ(1, i).re = cos(pz); EZ(1, i).im = sin(pz); EZ(-1, i).re = cos(pz); EZ(-1, i).im = -sin(pz); }
  }
  for (int j = 2; j <= kx; j++)
    for (int i = 0; i < ns; i++) {
      EX(j, i).re = EX(j - 1, i).re * EX(1, i).re - EX(j - 1, i).im * EX(1, i).im;
      EX(j, i).im = EX(j - 1, i).im * EX(1, i).re + EX(j - 1, i).re * EX(1, i).im;
    }
  for (int j = 2; j <= ky; j++)
    for (int i = 0; i < ns; i++) {
      EY(j, i).re = EY(j - 1, i).re * EY(1, i).re - EY(j - 1, i).im * EY(1, i).im;
      EY(j, i).im = EY(j - 1, i).im * EY(1, i).re + EY(j - 1, i).re * EY(1, i).im;
      EY(-j, i).re = EY(j, i).re;
      EY(-j, i).im = -EY(j, i).im;
    }
  for (int j = 2; j <= kz; j++)
    for (int i = 0; i < ns; i++) {
      EZ(j, i).re = EZ(j - 1, i).re * EZ(1, i).re - EZ(j - 1, i).im * EZ(1, i).im;
      EZ(j, i).im = EZ(j - 1, i).im * EZ(1, i).re + EZ(j - 1, i).re * EZ(1, i).im;
      EZ(-j, i).re = EZ(j, i).re;
      EZ(-j, i).im = -EZ(j, i).im;
    }
}
static void eik_free(eik_t *e) { free(e->ex); free(e->ey); free(e->ez); }

/* ConvertFromXYZtoABC then *TWO_PI (src/ewald.c:6990-6992, src/potentials.c:5566) */
static void frac_2pi(const orc_system *s, const double p[3], double out[3]) {
  const double *inv = s->inv;
  out[0] = inv[AX] * p[0] + inv[BX] * p[1] + inv[CX] * p[2];
  out[1] = inv[AY] * p[0] + inv[BY] * p[1] + inv[CY] * p[2];
  out[2] = inv[AZ] * p[0] + inv[BZ] * p[1] + inv[CZ] * p[2];
  out[0] *= (2.0 * M_PI);
  out[1] *= (2.0 * M_PI);
  out[2] *= (2.0 * M_PI);
}

/* Structure factors sum_i q_i exp(i k.r_i) over the k-vector list, in the
 * reference's loop order (src/ewald.c:1289-1731 framework part). */
static void structure_factors(const orc_system *s, int ns, const double *qs, const double (*fpos)[3],
                              cplx *out) {
  eik_t E;
  eik_t *e = &E;
  eik_build(e, ns, fpos, s->kx, s->ky, s->kz);
  cplx *xy = (cplx *)malloc(sizeof(cplx) * (size_t)(ns > 0 ? ns : 1));
  int kx = s->kx, ky = s->ky, kz = s->kz;
  int nvec = 0;
  for (int ii = 0; ii <= kx; ii++)
    for (int jj = -ky; jj <= ky; jj++) {
      for (int i = 0; i < ns; i++) {
        xy[i].re = EX(ii, i).re * EY(jj, i).re - EX(ii, i).im * EY(jj, i).im;
        xy[i].im = EX(ii, i).im * EY(jj, i).re + EX(ii, i).re * EY(jj, i).im;
      }
      for (int kk = -kz; kk <= kz; kk++) {
        double ksqr = SQR(ii) + SQR(jj) + SQR(kk);
        if ((ksqr != 0) && (ksqr < s->recip_cutoff)) {
          cplx sum = {0.0, 0.0};
          for (int i = 0; i < ns; i++) {
            double re = xy[i].re * EZ(kk, i).re - xy[i].im * EZ(kk, i).im;
            double im = xy[i].im * EZ(kk, i).re + xy[i].re * EZ(kk, i).im;
            double temp = qs[i];
            sum.re += temp * re;
            sum.im += temp * im;
          }
          out[nvec++] = sum;
        }
      }
    }
  free(xy);
  eik_free(e);
}
#undef EX
#undef EY
#undef EZ

/* EwaldEnergyIon (src/ewald.c:615-742): single +1 charge at the origin */
static void ewald_energy_ion(orc_system *s) {
  double energy = 0;
  double f[1][3] = {{0.0, 0.0, 0.0}};
  double one = 1.0;
  cplx *sk = (cplx *)malloc(sizeof(cplx) * (size_t)s->nk);
  structure_factors(s, 1, &one, (const double(*)[3])f, sk);
  for (int k = 0; k < s->nk; k++) energy += s->kfactor[k] * ((SQR(sk[k].re) + SQR(sk[k].im)));
  s->uion = -(energy - s->u.C * s->alpha / sqrt(M_PI));
  free(sk);
}

orc_system *orc_system_create(const double box[9], int bc, int n_atoms, const double *x,
                              const double *y, const double *z, const double *charge,
                              const double *eps_k, const double *sigma, const double *mass,
                              double cutoff_vdw, int shifted, int charge_method,
                              double cutoff_charge, double ewald_precision, double temperature) {
  orc_system *s = (orc_system *)calloc(1, sizeof(orc_system));
  double det;
  memcpy(s->box, box, sizeof(double) * 9);
  orc_invert3x3(s->box, s->inv, &det);
  orc_cell_properties(s->box, s->props);
  s->bc = bc;
  s->n = n_atoms;
  size_t nb = sizeof(double) * (size_t)(n_atoms > 0 ? n_atoms : 1);
  s->x = (double *)malloc(nb);
  s->y = (double *)malloc(nb);
  s->z = (double *)malloc(nb);
  s->q = (double *)malloc(nb);
  s->eps = (double *)malloc(nb);
  s->sig = (double *)malloc(nb);
  s->mass_total = 0.0;
  for (int i = 0; i < n_atoms; i++) {
    s->x[i] = x[i];
    s->y[i] = y[i];
    s->z[i] = z[i];
    s->q[i] = charge ? charge[i] : 0.0;
    s->eps[i] = eps_k[i];
    s->sig[i] = sigma[i];
    s->mass_total += mass ? mass[i] : 12.0;
  }
  s->u = make_units();
  s->rc_vdw2 = SQR(cutoff_vdw);
  s->shifted = shifted;
  s->charge_method = charge_method;
  s->rc_charge = cutoff_charge;
  s->rc_charge2 = SQR(cutoff_charge);
  s->temperature = temperature;
  s->beta = 1.0 / (s->u.KB * temperature); /* src/input.c:2071 */
  if (charge_method == ORC_CHARGE_EWALD) {
    ewald_init(s, ewald_precision);
    ewald_setup_kvectors(s);
    ewald_energy_ion(s);
    /* framework structure factor, charged atoms only (src/ewald.c:1330-1345) */
    int nc = 0;
    double(*fp)[3] = (double(*)[3])malloc(sizeof(double) * 3 * (size_t)(n_atoms > 0 ? n_atoms : 1));
    double *qc = (double *)malloc(nb);
    s->net_charge_framework = 0.0;
    for (int i = 0; i < n_atoms; i++) {
      if (considered_charged(s->q[i])) {
        double p[3] = {s->x[i], s->y[i], s->z[i]};
        qc[nc] = s->q[i];
        s->net_charge_framework += s->q[i];
        frac_2pi(s, p, fp[nc]);
        nc++;
      }
    }
    s->sf = (cplx *)malloc(sizeof(cplx) * (size_t)(s->nk > 0 ? s->nk : 1));
    structure_factors(s, nc, qc, (const double(*)[3])fp, s->sf);
    free(fp);
    free(qc);
  }
  return s;
}

void orc_system_destroy(orc_system *s) {
  if (s) {
    free(s->ads_start);
    free(s->ads_type);
    free(s->ads_mol);
    free(s->ads_xyz);
    free(s->sads);
  }
  if (s) free(s->pk);
  if (!s) return;
  free(s->x); free(s->y); free(s->z); free(s->q); free(s->eps); free(s->sig);
  for (int t = 0; t < s->ntypes; t++) free(s->pp[t]);
  free(s->kvec); free(s->kfactor); free(s->kidx); free(s->sf);
  free(s);
}

void orc_system_info(const orc_system *s, double out[8]) {
  out[0] = s->alpha;
  out[1] = s->kx;
  out[2] = s->ky;
  out[3] = s->kz;
  out[4] = s->nk;
  out[5] = s->uion;
  out[6] = s->beta;
  out[7] = s->props[6];
}

/* ===================================================================== */
/* energies at a position (src/framework_energy.c:60-311)                 */
/* ===================================================================== */
double orc_vdw_energy_at(const orc_system *s, int type, const double pos[3], double scaling) {
  (void)scaling; /* the LENNARD_JONES branch does not use it */
  /* grid branch (src/framework_energy.c:73-76) */
  if (s->vgrid[type]) return orc_interpolate_grid(s->vgrid[type], s->vgeom[type], s->bc, s->g_ubox, s->g_uinv, pos);
  double UVDW = 0.0;
  const double *pp = s->pp[type];
  for (int i = 0; i < s->n; i++) {
    double dr[3] = {pos[0] - s->x[i], pos[1] - s->y[i], pos[2] - s->z[i]};
    orc_min_image(s->box, s->inv, s->bc, dr);
    double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
    if (rr < s->rc_vdw2) UVDW += lj_value(pp[3 * i], pp[3 * i + 1], pp[3 * i + 2], rr);
  }
  return UVDW;
}

double orc_charge_energy_at(const orc_system *s, int type, const double pos[3], double scaling) {
  double U = 0.0;
  if (s->charge_method == ORC_CHARGE_NONE) return 0.0;
  if (!considered_charged(s->t_q[type])) return 0.0;
  /* grid branch (src/framework_energy.c:197-200): the charge of the
   * pseudo-atom, not scaled */
  if (s->cgrid) return orc_interpolate_coulomb_grid(s->cgrid, s->cgeom, s->bc, s->g_ubox, s->g_uinv, pos, s->t_q[type]);
  double chargeA = scaling * s->t_q[type];
  for (int i = 0; i < s->n; i++) {
    double dr[3] = {pos[0] - s->x[i], pos[1] - s->y[i], pos[2] - s->z[i]};
    orc_min_image(s->box, s->inv, s->bc, dr);
    double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
    if (rr < s->rc_charge2) {
      double r = sqrt(rr);
      double chargeB = s->q[i];
      /* PotentialValueCoulombic EWALD (src/potentials.c:12471-12472) */
      U += s->u.C * chargeA * chargeB * (erfc(s->alpha * r) / r);
    }
  }
  return U;
}

void orc_energy_at_many(const orc_system *s, int type, int64_t n, const double *xyz,
                        double *u_vdw, double *u_cc) {
  for (int64_t i = 0; i < n; i++) {
    if (u_vdw) u_vdw[i] = orc_vdw_energy_at(s, type, xyz + 3 * i, 1.0);
    if (u_cc) u_cc[i] = orc_charge_energy_at(s, type, xyz + 3 * i, 1.0);
  }
}

int orc_use_grid(orc_system *s, int type, const float *grid, const double geom[12], const double unit_box[9]) {
  if (type < 0 || type >= s->ntypes) return -1;
  s->vgrid[type] = grid;
  if (grid) {
    memcpy(s->vgeom[type], geom, 12 * sizeof(double));
    double det;
    memcpy(s->g_ubox, unit_box, 9 * sizeof(double));
    orc_invert3x3(s->g_ubox, s->g_uinv, &det);
  }
  return 0;
}

int orc_use_coulomb_grid(orc_system *s, const float *grid, const double geom[12], const double unit_box[9]) {
  s->cgrid = grid;
  if (grid) {
    memcpy(s->cgeom, geom, 12 * sizeof(double));
    double det;
    memcpy(s->g_ubox, unit_box, 9 * sizeof(double));
    orc_invert3x3(s->g_ubox, s->g_uinv, &det);
  }
  return 0;
}

/* ===================================================================== */
/* Ewald helpers                                                          */
/* ===================================================================== */
int orc_ewald_nk(const orc_system *s) { return s->nk; }

void orc_ewald_tables(const orc_system *s, double *kvec3, double *kfactor, double *sf_re,
                      double *sf_im, int *kidx3) {
  for (int k = 0; k < s->nk; k++) {
    if (kvec3) {
      kvec3[3 * k] = s->kvec[3 * k];
      kvec3[3 * k + 1] = s->kvec[3 * k + 1];
      kvec3[3 * k + 2] = s->kvec[3 * k + 2];
    }
    if (kidx3) {
      kidx3[3 * k] = s->kidx[3 * k];
      kidx3[3 * k + 1] = s->kidx[3 * k + 1];
      kidx3[3 * k + 2] = s->kidx[3 * k + 2];
    }
    if (kfactor) kfactor[k] = s->kfactor[k];
    if (sf_re) sf_re[k] = s->sf[k].re;
    if (sf_im) sf_im[k] = s->sf[k].im;
  }
}

double orc_ewald_fourier_self_energy_framework(const orc_system *s) {
  double e = 0.0;
  for (int k = 0; k < s->nk; k++) e += s->kfactor[k] * (SQR(s->sf[k].re) + SQR(s->sf[k].im));
  return e;
}

double orc_ewald_self_energy_framework(const orc_system *s) {
  double e = 0.0;
  for (int i = 0; i < s->n; i++)
    if (considered_charged(s->q[i])) e += s->u.C * SQR(s->q[i]) * s->alpha / sqrt(M_PI);
  return e;
}

double orc_ewald_real_energy_framework(const orc_system *s) {
  double e = 0.0;
  for (int i = 0; i < s->n; i++)
    for (int j = i + 1; j < s->n; j++) {
      double dr[3] = {s->x[i] - s->x[j], s->y[i] - s->y[j], s->z[i] - s->z[j]};
      orc_min_image(s->box, s->inv, s->bc, dr);
      double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
      if (rr < s->rc_charge2) {
        double r = sqrt(rr);
        e += s->u.C * s->q[i] * s->q[j] * (erfc(s->alpha * r) / r);
      }
    }
  return e;
}

/* CalculateEwaldFourierAdsorbate(TRUE,FALSE,mol,0) at zero loading
 * (src/ewald.c:6859-7565); no bond dipoles, no cations. */
void orc_fourier_move(const orc_system *s, int n_new, const int *new_types, const double *new_xyz,
                      int old_mol, double out[5]) {
  double charge[2 * ORC_MAX_SITES], fpos[2 * ORC_MAX_SITES][3], cart[2 * ORC_MAX_SITES][3];
  int nc = 0;
  double energy_self_new = 0.0, energy_self_old = 0.0, net_charge_new = 0.0, net_charge_old = 0.0;
  double alpha = s->alpha;
  for (int i = 0; i < 5; i++) out[i] = 0.0;
  if (s->charge_method != ORC_CHARGE_EWALD) return;
  /* old molecule's charged atoms first (src/ewald.c:6944-6962) */
  if (old_mol >= 0) {
    for (int a = s->ads_start[old_mol]; a < s->ads_start[old_mol + 1]; a++) {
      double q = s->t_q[s->ads_type[a]];
      if (considered_charged(q)) {
        charge[nc] = q * 1.0;
        energy_self_old += s->u.C * SQR(charge[nc]) * alpha / sqrt(M_PI);
        net_charge_old += charge[nc];
        const double *p = s->ads_xyz + 3 * a;
        frac_2pi(s, p, fpos[nc]);
        cart[nc][0] = p[0];
        cart[nc][1] = p[1];
        cart[nc][2] = p[2];
        nc++;
      }
    }
  }
  const int nc_old = nc;
  /* then the new molecule's (src/ewald.c:6965-6981) */
  for (int i = 0; i < n_new; i++) {
    int type = new_types[i];
    if (considered_charged(s->t_q[type])) {
      double scaling = 1.0;
      charge[nc] = scaling * s->t_q[type];
      energy_self_new += s->u.C * SQR(charge[nc]) * alpha / sqrt(M_PI);
      net_charge_new += charge[nc];
      double p[3] = {new_xyz[3 * i], new_xyz[3 * i + 1], new_xyz[3 * i + 2]};
      frac_2pi(s, p, fpos[nc]);
      cart[nc][0] = p[0];
      cart[nc][1] = p[1];
      cart[nc][2] = p[2];
      nc++;
    }
  }
  const int nc_new = nc - nc_old;
  size_t nk = (size_t)(s->nk > 0 ? s->nk : 1);
  cplx *so = (cplx *)malloc(sizeof(cplx) * nk), *sn = (cplx *)malloc(sizeof(cplx) * nk);
  structure_factors(s, nc_old, charge, (const double(*)[3])fpos, so);
  structure_factors(s, nc_new, charge + nc_old, (const double(*)[3])(fpos + nc_old), sn);
  double energy_charge_framework_adsorbates = 0.0, energy_charge_adsorbates = 0.0;
  for (int k = 0; k < s->nk; k++) {
    /* src/ewald.c:7233-7259 */
    cplx st = s->sads ? s->sads[k] : (cplx){0.0, 0.0};
    cplx sa;
    sa.re = st.re + (sn[k].re - so[k].re);
    sa.im = st.im + (sn[k].im - so[k].im);
    double temp = s->kfactor[k];
    energy_charge_framework_adsorbates += temp * (s->sf[k].re * (sa.re - st.re) + s->sf[k].im * (sa.im - st.im));
    energy_charge_adsorbates += temp * (SQR(sa.re) - SQR(st.re) + SQR(sa.im) - SQR(st.im));
  }
  free(so);
  free(sn);
  /* excluded intra pairs: all charged pairs i<j of each molecule
   * (src/molecule.c:2839-2851, src/ewald.c:7318-7335, 7418-7440) */
  double energy_excluded_new = 0.0, energy_excluded_old = 0.0;
  for (int part = 0; part < 2; part++) {
    int b0 = part ? nc_old : 0, b1 = part ? nc : nc_old;
    double e = 0.0;
    for (int a = b0; a < b1; a++)
      for (int b = a + 1; b < b1; b++) {
        double dx = cart[a][0] - cart[b][0];
        double dy = cart[a][1] - cart[b][1];
        double dz = cart[a][2] - cart[b][2];
        double r = sqrt(SQR(dx) + SQR(dy) + SQR(dz));
        double Bt0 = -erf(alpha * r) / r;
        e -= s->u.C * charge[a] * charge[b] * Bt0;
      }
    if (part) energy_excluded_new = e;
    else energy_excluded_old = e;
  }
  /* src/ewald.c:7536-7565 */
  double net_delta = net_charge_new - net_charge_old;
  double nca = s->net_charge_ads;
  double ads = energy_charge_adsorbates - (energy_excluded_new - energy_excluded_old) -
               (energy_self_new - energy_self_old);
  ads += s->uion * SQR(nca + net_delta) - s->uion * SQR(nca);
  double host = 2.0 * energy_charge_framework_adsorbates;
  host += 2.0 * s->uion * s->net_charge_framework * (nca + net_delta) -
          2.0 * s->uion * s->net_charge_framework * (nca);
  out[0] = host;
  out[1] = ads;
  out[2] = energy_self_new - energy_self_old;
  out[3] = energy_excluded_new - energy_excluded_old;
  out[4] = energy_charge_adsorbates;
}

void orc_fourier_adsorbate(const orc_system *s, int n_sites, const int *types,
                           const double *xyz, double out[5]) {
  orc_fourier_move(s, n_sites, types, xyz, -1, out);
}

/* SurfaceAreaMove / CheckSurfaceAreaOverlap (src/mc_moves.c:16434-16520,
 * src/framework.c:11649-11720), with the points of the counter stream of
 * (seed, atom x samples + k) and RandomNumberOnUnitSphere's rejection
 * (src/utils.c:688-708). */
double orc_surface_area(const orc_system *s, int probe_type, double factor, int64_t samples, uint64_t seed,
                        const uint8_t *interacts, uint64_t *counts) {
  double area = 0.0;
  const double sp = s->t_sig[probe_type];
  for (int i = 0; i < s->n; i++) {
    /* only atoms whose pseudo-atom interacts carry a sphere (src/mc_moves.c:16456) */
    if (interacts && !interacts[i]) {
      if (counts) counts[i] = 0;
      continue;
    }
    double di = factor * (0.5 * (sp + s->sig[i]));
    double total = 0.0, counted = 0.0;
    uint64_t cnt = 0;
    for (int64_t k = 0; k < samples; k++) {
      total += 1.0;
      uint64_t key = orc_stream_key(seed, (uint64_t)i * (uint64_t)samples + (uint64_t)k);
      uint64_t c = 0;
      double ran1, ran2, ransq;
      do {
        ran1 = 2.0 * orc_stream_uniform(key, c++) - 1.0;
        ran2 = 2.0 * orc_stream_uniform(key, c++) - 1.0;
        ransq = SQR(ran1) + SQR(ran2);
      } while (ransq >= 1.0);
      double ranh = 2.0 * sqrt(1.0 - ransq);
      double v[3] = {ran1 * ranh, ran2 * ranh, 1.0 - 2.0 * ransq};
      double p[3] = {s->x[i] + v[0] * di, s->y[i] + v[1] * di, s->z[i] + v[2] * di};
      if (orc_blocked_pocket(s, p)) continue;
      int overlap = 0;
      for (int j = 0; j < s->n && !overlap; j++) {
        if (j == i) continue;
        /* a NONE pair has sigma 0 in PotentialParms: rr < 0 never holds
           (src/framework.c:11676) */
        double dj = (interacts && !interacts[j]) ? 0.0 : factor * (0.5 * (sp + s->sig[j]));
        double dr[3] = {p[0] - s->x[j], p[1] - s->y[j], p[2] - s->z[j]};
        orc_min_image(s->box, s->inv, s->bc, dr);
        double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
        if (rr < SQR(dj)) overlap = 1;
      }
      if (!overlap) {
        counted += 1.0;
        cnt++;
      }
    }
    if (counts) counts[i] = cnt;
    area += (counted / total) * 4.0 * M_PI * SQR(di);
  }
  return area;
}

/* LB-mixed LJ parameters of the pair (typeA, typeB) as orc_add_pseudo_atom
 * forms them for framework atoms: eps (internal units), sigma, shift */
static void pair_lj(const orc_system *s, int ta, int tb, double out[3]);

int orc_set_adsorbates(orc_system *s, int n_mol, const int *start, const int *types, const double *xyz) {
  free(s->ads_start);
  free(s->ads_type);
  free(s->ads_mol);
  free(s->ads_xyz);
  free(s->sads);
  s->ads_start = s->ads_type = s->ads_mol = NULL;
  s->ads_xyz = NULL;
  s->sads = NULL;
  s->ads_nmol = s->ads_n = 0;
  s->net_charge_ads = 0.0;
  if (n_mol <= 0) return 0;
  int n = start[n_mol];
  for (int i = 0; i < n; i++)
    if (types[i] < 0 || types[i] >= s->ntypes) return -1;
  s->ads_nmol = n_mol;
  s->ads_n = n;
  s->ads_start = (int *)malloc(sizeof(int) * (size_t)(n_mol + 1));
  s->ads_type = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  s->ads_mol = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  s->ads_xyz = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  memcpy(s->ads_start, start, sizeof(int) * (size_t)(n_mol + 1));
  memcpy(s->ads_type, types, sizeof(int) * (size_t)n);
  memcpy(s->ads_xyz, xyz, sizeof(double) * 3 * (size_t)n);
  for (int m = 0; m < n_mol; m++)
    for (int a = start[m]; a < start[m + 1]; a++) s->ads_mol[a] = m;
  /* StoreTotalChargeAdsorbates over every charged adsorbate atom in order */
  double *q = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double(*fp)[3] = (double(*)[3])malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  int nc = 0;
  for (int a = 0; a < n; a++) {
    double qa = s->t_q[types[a]];
    if (considered_charged(qa)) {
      q[nc] = qa;
      s->net_charge_ads += qa;
      frac_2pi(s, xyz + 3 * a, fp[nc]);
      nc++;
    }
  }
  if (s->charge_method == ORC_CHARGE_EWALD) {
    s->sads = (cplx *)malloc(sizeof(cplx) * (size_t)(s->nk > 0 ? s->nk : 1));
    structure_factors(s, nc, q, (const double(*)[3])fp, s->sads);
  }
  free(q);
  free(fp);
  return 0;
}

double orc_inter_vdw_at(const orc_system *s, int type, const double pos[3], int exclude, double scaling) {
  (void)scaling; /* the LENNARD_JONES branch does not use it */
  double Uvdw = 0.0;
  for (int j = 0; j < s->ads_nmol; j++) {
    if (j == exclude) continue;
    for (int a = s->ads_start[j]; a < s->ads_start[j + 1]; a++) {
      double dr[3] = {pos[0] - s->ads_xyz[3 * a], pos[1] - s->ads_xyz[3 * a + 1], pos[2] - s->ads_xyz[3 * a + 2]};
      orc_min_image(s->box, s->inv, s->bc, dr);
      double r2 = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
      if (r2 < s->rc_vdw2) {
        double pp[3];
        pair_lj(s, type, s->ads_type[a], pp);
        Uvdw += lj_value(pp[0], pp[1], pp[2], r2);
      }
    }
  }
  return Uvdw;
}

double orc_inter_charge_at(const orc_system *s, int type, const double pos[3], int exclude, double charge_a) {
  (void)type;
  double U = 0.0;
  if (s->charge_method == ORC_CHARGE_NONE) return 0.0; /* PotentialValueCoulombic NONE */
  for (int j = 0; j < s->ads_nmol; j++) {
    if (j == exclude) continue;
    for (int a = s->ads_start[j]; a < s->ads_start[j + 1]; a++) {
      double chargeB = 1.0 * s->t_q[s->ads_type[a]];
      double dr[3] = {pos[0] - s->ads_xyz[3 * a], pos[1] - s->ads_xyz[3 * a + 1], pos[2] - s->ads_xyz[3 * a + 2]};
      orc_min_image(s->box, s->inv, s->bc, dr);
      double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
      if (rr < s->rc_charge2) {
        double r = sqrt(rr);
        U += s->u.C * charge_a * chargeB * (erfc(s->alpha * r) / r);
      }
    }
  }
  return U;
}

/* ===================================================================== */
/* Widom / CBMC (src/cbmc.c:521-795,1154-1578,2762-3185,                  */
/*               src/mc_moves.c:7844-7946, src/utils.c:688-877)           */
/* ===================================================================== */
typedef struct {
  orc_uniform_fn next;
  void *state;
  const orc_trials *given;
  orc_trials *record;
  int portable_trig; /* Shoemake trig from sincos_portable (counter-stream trials) */
} draw_t;

static double draw(draw_t *d) { return d->next(d->state); }

/* SelectTrialPosition (src/cbmc.c:563-595) */
static int select_trial(const double *bf, const int *overlap, int n, double u) {
  double largest = -DBL_MAX, sum = 0.0, shifted[64];
  for (int i = 0; i < n; i++)
    if (!overlap[i])
      if (bf[i] > largest) largest = bf[i];
  for (int i = 0; i < n; i++) {
    if (!overlap[i]) {
      shifted[i] = exp(bf[i] - largest);
      sum += shifted[i];
    } else
      shifted[i] = 0.0;
  }
  int selected = 0;
  double cumw = shifted[0];
  double ws = u * sum;
  while (cumw < ws) cumw += shifted[++selected];
  return selected;
}

/* RandomNumberOnUnitSphere (src/utils.c:688-708); the vector is unused for rigid groups */
static void unit_sphere(draw_t *d) {
  double ran1, ran2, ransq;
  do {
    ran1 = 2.0 * draw(d) - 1.0;
    ran2 = 2.0 * draw(d) - 1.0;
    ransq = SQR(ran1) + SQR(ran2);
  } while (ransq >= 1.0);
}

/* RandomArrayRotationMatrix (src/utils.c:835-877) */
static void random_rotation(draw_t *d, double (*cord)[3], int n) {
  double R[3][3];
  double X0 = draw(d);
  double Y1 = 2.0 * M_PI * draw(d);
  double Y2 = 2.0 * M_PI * draw(d);
  double R1 = sqrt(1.0 - X0);
  double R2 = sqrt(X0);
  double sY1, cY1, sY2, cY2;
  if (d->portable_trig) {
    sincos_portable(Y1, &sY1, &cY1);
    sincos_portable(Y2, &sY2, &cY2);
  } else {
    sY1 = sin(Y1);
    cY1 = cos(Y1);
    sY2 = sin(Y2);
    cY2 = cos(Y2);
  }
  double U0 = cY2 * R2;
  double U1 = sY1 * R1;
  double U2 = cY1 * R1;
  double U3 = sY2 * R2;
  double COEFI = 2.0 * U0 * U0 - 1.0;
  double COEFUU = 2.0;
  double COEFE = 2.0 * U0;
  R[0][0] = COEFI + COEFUU * U1 * U1;
  R[1][1] = COEFI + COEFUU * U2 * U2;
  R[2][2] = COEFI + COEFUU * U3 * U3;
  R[1][2] = COEFUU * U2 * U3 - COEFE * U1;
  R[2][0] = COEFUU * U3 * U1 - COEFE * U2;
  R[0][1] = COEFUU * U1 * U2 - COEFE * U3;
  R[2][1] = COEFUU * U3 * U2 + COEFE * U1;
  R[0][2] = COEFUU * U1 * U3 + COEFE * U2;
  R[1][0] = COEFUU * U2 * U1 + COEFE * U3;
  for (int i = 0; i < n; i++) {
    double px = cord[i][0] * R[0][0] + cord[i][1] * R[0][1] + cord[i][2] * R[0][2];
    double py = cord[i][0] * R[1][0] + cord[i][1] * R[1][1] + cord[i][2] * R[1][2];
    double pz = cord[i][0] * R[2][0] + cord[i][1] * R[2][1] + cord[i][2] * R[2][2];
    cord[i][0] = px;
    cord[i][1] = py;
    cord[i][2] = pz;
  }
}

/* retrace of an existing molecule (RetraceMolecule with CBMC_DELETION,
 * src/cbmc.c:2928-3054, 623-795, 2851-2924): trial 0 of both stages is the
 * molecule's own configuration `old` (n_sites x 3) and is always "chosen",
 * the molecule `exclude` is left out of the guest-guest sums, and there is no
 * minimum-Rosenbluth, blocked-pocket or Fourier step */
typedef struct {
  int retrace;
  int exclude;
  const double *old;
} core_opts;

static void widom_core_opts(const orc_system *s, const orc_molecule *m, draw_t *d, double rec[ORC_REC_N],
                            const core_opts *o);

static void widom_core(const orc_system *s, const orc_molecule *m, draw_t *d, double rec[ORC_REC_N]) {
  core_opts o = {0, s->ads_nmol, NULL};
  widom_core_opts(s, m, d, rec, &o);
}

static void widom_core_opts(const orc_system *s, const orc_molecule *m, draw_t *d, double rec[ORC_REC_N],
                            const core_opts *opt) {
  const double crit = ENERGY_OVERLAP_CRITERIA;
  const double beta = s->beta;
  int k_fb = m->k_fb, k_or = m->k_or;
  int start = m->start;
  double bf[64], evdw[64], ecc[64], eav[64], eac[64];
  int ov[64];
  /* the inserted molecule is adsorbate number ads_nmol: nothing excluded
   * (src/mc_moves.c:7856); a retraced one excludes itself */
  const int excl = opt->exclude;
  const int retrace = opt->retrace;
  double fbpos[64][3];
  for (int i = 0; i < ORC_REC_N; i++) rec[i] = 0.0;

  /* HandleFirstBead(CBMC_INSERTION) (src/cbmc.c:623-795) */
  int OVERLAP = 1;
  uint64_t fbmask = 0;
  int type0 = m->type[start];
  for (int i = 0; i < k_fb; i++) {
    double sv[3];
    if (retrace && i == 0) {
      /* FirstBeadPosition = OldPosition[start] (src/cbmc.c:2990) */
      fbpos[0][0] = opt->old[3 * start];
      fbpos[0][1] = opt->old[3 * start + 1];
      fbpos[0][2] = opt->old[3 * start + 2];
    } else if (d->given) {
      fbpos[i][0] = d->given->fb[i][0];
      fbpos[i][1] = d->given->fb[i][1];
      fbpos[i][2] = d->given->fb[i][2];
    } else {
      sv[2] = draw(d); /* fall-through z, y, x (src/cbmc.c:652-667) */
      sv[1] = draw(d);
      sv[0] = draw(d);
      /* ConvertFromABCtoXYZ (src/potentials.c:5548-5556) */
      fbpos[i][0] = s->box[AX] * sv[0] + s->box[BX] * sv[1] + s->box[CX] * sv[2];
      fbpos[i][1] = s->box[AY] * sv[0] + s->box[BY] * sv[1] + s->box[CY] * sv[2];
      fbpos[i][2] = s->box[AZ] * sv[0] + s->box[BZ] * sv[1] + s->box[CZ] * sv[2];
    }
    if (d->record) {
      d->record->fb[i][0] = fbpos[i][0];
      d->record->fb[i][1] = fbpos[i][1];
      d->record->fb[i][2] = fbpos[i][2];
    }
    double Uvdw = orc_vdw_energy_at(s, type0, fbpos[i], 1.0);
    double Ucc = orc_charge_energy_at(s, type0, fbpos[i], 1.0);
    /* guest-guest terms (src/cbmc.c:683-689) */
    double Uav = orc_inter_vdw_at(s, type0, fbpos[i], excl, 1.0);
    double Uac = orc_inter_charge_at(s, type0, fbpos[i], excl, 1.0 * s->t_q[type0]);
    if ((Uvdw >= crit) || (Ucc >= crit) || (Uav >= crit) || (Uac >= crit)) {
      ov[i] = 1;
      fbmask |= (1ULL << i);
    } else {
      ov[i] = 0;
      OVERLAP = 0;
      evdw[i] = Uvdw;
      ecc[i] = Ucc;
      eav[i] = Uav;
      eac[i] = Uac;
      bf[i] = -beta * (Uvdw + Uav + 0.0);
    }
  }
  rec[ORC_REC_FBMASK] = (double)fbmask;
  /* a retraced configuration that itself overlaps is reported as overlap
   * (the reference would read stale trial energies; accepted moves never
   * produce such a configuration) */
  if (retrace && ov[0]) OVERLAP = 1;
  if (OVERLAP) {
    rec[ORC_REC_OVERLAP] = 1.0;
    rec[ORC_REC_STAGE] = 0.0;
    return;
  }
  double W1 = 0.0;
  for (int i = 0; i < k_fb; i++)
    if (!ov[i]) W1 += exp(bf[i]);
  int sel = 0;
  if (!retrace) {
    double usel = d->given ? d->given->u_sel_fb : draw(d);
    if (d->record) d->record->u_sel_fb = usel;
    sel = select_trial(bf, ov, k_fb, usel);
  }
  W1 /= (double)k_fb;
  rec[ORC_REC_SELFB] = sel;
  rec[ORC_REC_W1] = W1;
  if (!retrace && W1 < MINIMUM_ROSENBLUTH_FACTOR) {
    rec[ORC_REC_OVERLAP] = 1.0;
    rec[ORC_REC_STAGE] = 1.0;
    return;
  }
  double RosenbluthNew = W1;
  double UHostVDWNew = evdw[sel];
  double UHostChargeChargeNew = ecc[sel];
  double UAdsorbateVDWNew = eav[sel];
  double UAdsorbateChargeChargeNew = eac[sel];
  double newpos[ORC_MAX_SITES][3];
  newpos[start][0] = fbpos[sel][0];
  newpos[start][1] = fbpos[sel][1];
  newpos[start][2] = fbpos[sel][2];

  if (m->n_sites > 1) {
    /* Rosen -> GenerateTrialOrientationsSimpleSphere -> ComputeExternalEnergies */
    double tpos[64][ORC_MAX_SITES][3];
    double uv[64], uc[64], ua[64], uac[64];
    int ovo[64];
    uint64_t ormask = 0;
    int n = m->n_sites;
    for (int iu = 0; iu < k_or; iu++) {
      double cord[ORC_MAX_SITES][3];
      if (d->given) {
        for (int j = 0; j < n; j++) {
          cord[j][0] = d->given->cord[iu][j][0];
          cord[j][1] = d->given->cord[iu][j][1];
          cord[j][2] = d->given->cord[iu][j][2];
        }
      } else {
        unit_sphere(d);
        for (int j = 0; j < n; j++) {
          cord[j][0] = m->body[j][0] - m->body[start][0];
          cord[j][1] = m->body[j][1] - m->body[start][1];
          cord[j][2] = m->body[j][2] - m->body[start][2];
        }
        random_rotation(d, cord, n);
      }
      if (d->record)
        for (int j = 0; j < n; j++) {
          d->record->cord[iu][j][0] = cord[j][0];
          d->record->cord[iu][j][1] = cord[j][1];
          d->record->cord[iu][j][2] = cord[j][2];
        }
      for (int j = 0; j < n; j++) {
        if (retrace && iu == 0) {
          /* TrialPositions[0] = OldPosition (src/cbmc.c:2862-2864) */
          tpos[0][j][0] = opt->old[3 * j];
          tpos[0][j][1] = opt->old[3 * j + 1];
          tpos[0][j][2] = opt->old[3 * j + 2];
          continue;
        }
        tpos[iu][j][0] = newpos[start][0] + cord[j][0];
        tpos[iu][j][1] = newpos[start][1] + cord[j][1];
        tpos[iu][j][2] = newpos[start][2] + cord[j][2];
      }
    }
    int ALLOV = 1;
    for (int iu = 0; iu < k_or; iu++) {
      double EV = 0.0, EC = 0.0, EA = 0.0, EAC = 0.0;
      int trial_overlap = 0;
      for (int j = 0; j < n; j++) {
        if (j == start) continue;
        double Uvdw = orc_vdw_energy_at(s, m->type[j], tpos[iu][j], 1.0);
        if (Uvdw >= crit) {
          trial_overlap = 1;
          break;
        }
        double Ucc = orc_charge_energy_at(s, m->type[j], tpos[iu][j], 1.0);
        EV += Uvdw;
        EC += Ucc;
        /* guest-guest terms: summed, not checked for overlap (src/cbmc.c:1222-1244) */
        double Uav = orc_inter_vdw_at(s, m->type[j], tpos[iu][j], excl, 1.0);
        double Uac = orc_inter_charge_at(s, m->type[j], tpos[iu][j], excl, 1.0 * s->t_q[m->type[j]]);
        EA += Uav;
        EAC += Uac;
      }
      if (!trial_overlap) {
        uv[iu] = EV;
        uc[iu] = EC;
        ua[iu] = EA;
        uac[iu] = EAC;
        bf[iu] = -beta * (EV + EA + 0.0 + 0.0);
        ovo[iu] = 0;
        ALLOV = 0;
      } else {
        bf[iu] = 0.0;
        ovo[iu] = 1;
        ormask |= (1ULL << iu);
      }
    }
    rec[ORC_REC_ORMASK] = (double)ormask;
    if (retrace && ovo[0]) ALLOV = 1;
    if (ALLOV) {
      rec[ORC_REC_OVERLAP] = 1.0;
      rec[ORC_REC_STAGE] = 2.0;
      return;
    }
    double weight = 0.0;
    for (int i = 0; i < k_or; i++)
      if (!ovo[i]) weight += exp(bf[i]);
    weight = weight / (double)k_or;
    int iwalk = 0;
    if (!retrace) {
      double uso = d->given ? d->given->u_sel_or : draw(d);
      if (d->record) d->record->u_sel_or = uso;
      iwalk = select_trial(bf, ovo, k_or, uso);
    }
    rec[ORC_REC_SELOR] = iwalk;
    rec[ORC_REC_W2] = weight;
    RosenbluthNew *= weight * 1.0;
    if (!retrace && RosenbluthNew < MINIMUM_ROSENBLUTH_FACTOR) {
      rec[ORC_REC_OVERLAP] = 1.0;
      rec[ORC_REC_STAGE] = 3.0;
      return;
    }
    /* src/cbmc.c:2819-2823 */
    UAdsorbateVDWNew += ua[iwalk];
    UHostVDWNew += uv[iwalk];
    UHostChargeChargeNew += uc[iwalk];
    UAdsorbateChargeChargeNew += uac[iwalk];
    for (int j = 0; j < n; j++)
      if (j != start) {
        newpos[j][0] = tpos[iwalk][j][0];
        newpos[j][1] = tpos[iwalk][j][1];
        newpos[j][2] = tpos[iwalk][j][2];
      }
  }
  /* GrowMolecule tail (src/cbmc.c:3133-3185): anisotropic, replica and
   * bond cross-term corrections are exactly zero here; LJ_BIASING
   * correction with the real-space Coulomb energy. */
  RosenbluthNew *= exp(-beta * (0.0 + 0.0 + 0.0));
  RosenbluthNew *= exp(-beta * (0.0 + 0.0));
  RosenbluthNew *= exp(-beta * (0.0 + 0.0));
  RosenbluthNew *= exp(-beta * (UHostChargeChargeNew + UAdsorbateChargeChargeNew + 0.0));
  /* WidomAdsorbateMove (src/mc_moves.c:7882-7886): a grown molecule with an
   * atom in a blocked pocket contributes nothing */
  for (int j = 0; j < m->n_sites && !retrace; j++)
    if (orc_blocked_pocket(s, newpos[j])) {
      rec[ORC_REC_W] = 0.0;
      rec[ORC_REC_DU] = 0.0;
      rec[ORC_REC_UVDW] = 0.0;
      rec[ORC_REC_UREAL] = 0.0;
      rec[ORC_REC_UFHOST] = 0.0;
      rec[ORC_REC_UFADS] = 0.0;
      rec[ORC_REC_OVERLAP] = 0.0;
      rec[ORC_REC_STAGE] = 5.0;
      return;
    }
  /* WidomAdsorbateMove (src/mc_moves.c:7878-7933): tail correction 0 */
  RosenbluthNew *= exp(-beta * 0.0);
  double fo[5] = {0, 0, 0, 0, 0};
  if (s->charge_method == ORC_CHARGE_EWALD && !retrace) {
    double xyz[3 * ORC_MAX_SITES];
    for (int j = 0; j < m->n_sites; j++) {
      xyz[3 * j] = newpos[j][0];
      xyz[3 * j + 1] = newpos[j][1];
      xyz[3 * j + 2] = newpos[j][2];
    }
    orc_fourier_adsorbate(s, m->n_sites, m->type, xyz, fo);
    RosenbluthNew *= exp(-beta * (fo[0] + fo[1] + 0.0));
  }
  /* src/mc_moves.c:7903-7913, the zero terms left out */
  double DeltaU = UAdsorbateVDWNew + UHostVDWNew + UAdsorbateChargeChargeNew + UHostChargeChargeNew + fo[0] + fo[1];
  rec[ORC_REC_W] = RosenbluthNew;
  rec[ORC_REC_DU] = DeltaU;
  rec[ORC_REC_UVDW] = UHostVDWNew;
  rec[ORC_REC_UREAL] = UHostChargeChargeNew;
  rec[ORC_REC_UFHOST] = fo[0];
  rec[ORC_REC_UFADS] = fo[1];
  rec[ORC_REC_UADSVDW] = UAdsorbateVDWNew;
  rec[ORC_REC_UADSREAL] = UAdsorbateChargeChargeNew;
  rec[ORC_REC_OVERLAP] = 0.0;
  rec[ORC_REC_STAGE] = 4.0;
}

static void insertion_impl(const orc_system *s, const orc_molecule *m, orc_uniform_fn next,
                           void *state, double rec[ORC_REC_N], orc_trials *trials_out, int portable) {
  draw_t d;
  memset(&d, 0, sizeof(d));
  d.next = next;
  d.state = state;
  d.record = trials_out;
  d.portable_trig = portable;
  widom_core(s, m, &d, rec);
}

void orc_widom_insertion(const orc_system *s, const orc_molecule *m, orc_uniform_fn next,
                         void *state, double rec[ORC_REC_N], orc_trials *trials_out) {
  insertion_impl(s, m, next, state, rec, trials_out, 0);
}

void orc_widom_retrace_trials(const orc_system *s, const orc_molecule *m, int exclude, const double *old,
                              const orc_trials *t, double rec[ORC_REC_N]) {
  draw_t d;
  memset(&d, 0, sizeof(d));
  d.given = t;
  core_opts o = {1, exclude < 0 ? s->ads_nmol : exclude, old};
  widom_core_opts(s, m, &d, rec, &o);
}

void orc_widom_insertion_trials(const orc_system *s, const orc_molecule *m,
                                const orc_trials *t, double rec[ORC_REC_N]) {
  draw_t d;
  memset(&d, 0, sizeof(d));
  d.given = t;
  widom_core(s, m, &d, rec);
}

typedef struct { uint64_t key; uint64_t ctr; } stream_state;
static double stream_next(void *p) {
  stream_state *st = (stream_state *)p;
  return orc_stream_uniform(st->key, st->ctr++);
}

/* Counter-stream layout of one insertion: 3*k_fb fb draws, one fb
 * selection draw, then the orientation draws, then one selection draw.
 * The fb selection draw index is fixed, so orientation draws start at
 * 3*k_fb+1 whatever the first-bead outcome. */
void orc_widom_run_stream(const orc_system *s, const orc_molecule *m, uint64_t seed, int64_t first,
                          int64_t n, double *rec) {
  for (int64_t i = 0; i < n; i++) {
    stream_state st = {orc_stream_key(seed, (uint64_t)(first + i)), 0};
    insertion_impl(s, m, stream_next, &st, rec + ORC_REC_N * i, NULL, 1);
  }
}

void orc_stream_trials(const orc_system *s, const orc_molecule *m, uint64_t seed, int64_t insertion,
                       double *fb, double *cord, double *u_sel) {
  /* replay the draw layout without energies */
  stream_state st = {orc_stream_key(seed, (uint64_t)insertion), 0};
  draw_t d;
  memset(&d, 0, sizeof(d));
  d.next = stream_next;
  d.state = &st;
  d.portable_trig = 1;
  for (int i = 0; i < m->k_fb; i++) {
    double sv[3];
    sv[2] = draw(&d);
    sv[1] = draw(&d);
    sv[0] = draw(&d);
    fb[3 * i + 0] = s->box[AX] * sv[0] + s->box[BX] * sv[1] + s->box[CX] * sv[2];
    fb[3 * i + 1] = s->box[AY] * sv[0] + s->box[BY] * sv[1] + s->box[CY] * sv[2];
    fb[3 * i + 2] = s->box[AZ] * sv[0] + s->box[BZ] * sv[1] + s->box[CZ] * sv[2];
  }
  u_sel[0] = draw(&d);
  u_sel[1] = 0.0;
  if (m->n_sites > 1) {
    int n = m->n_sites;
    for (int iu = 0; iu < m->k_or; iu++) {
      double c[ORC_MAX_SITES][3];
      unit_sphere(&d);
      for (int j = 0; j < n; j++) {
        c[j][0] = m->body[j][0] - m->body[m->start][0];
        c[j][1] = m->body[j][1] - m->body[m->start][1];
        c[j][2] = m->body[j][2] - m->body[m->start][2];
      }
      random_rotation(&d, c, n);
      for (int j = 0; j < n; j++) {
        cord[(iu * n + j) * 3 + 0] = c[j][0];
        cord[(iu * n + j) * 3 + 1] = c[j][1];
        cord[(iu * n + j) * 3 + 2] = c[j][2];
      }
    }
    u_sel[1] = draw(&d);
  }
}

/* block index per cycle (src/statistics.c:898-910, 1126-1127) */
void orc_block_of_cycles(int64_t n_cycles, int32_t *block) {
  long long BlockCycle[NR_BLOCKS];
  for (int i = 0; i < NR_BLOCKS; i++) BlockCycle[i] = (long long)((i + 1) * n_cycles / (double)NR_BLOCKS);
  int Block = 0;
  for (int64_t c = 0; c < n_cycles; c++) {
    if (Block < NR_BLOCKS && c == BlockCycle[Block]) Block++;
    block[c] = Block < NR_BLOCKS ? Block : NR_BLOCKS - 1;
  }
}

typedef struct { orc_mt64 g; } mt_state;
static double mt_next(void *p) { return orc_mt64_uniform(&((mt_state *)p)->g); }

int64_t orc_widom_run_mt_trials(const orc_system *s, const orc_molecule *m, uint64_t seed, int64_t n_cycles,
                                int steps_per_cycle, double *blocks, int64_t max_ins, double *recs,
                                double *fb, double *cord, double *u_sel) {
  mt_state st;
  orc_mt64_seed(&st.g, seed);
  int32_t *bmap = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n_cycles > 0 ? n_cycles : 1));
  orc_block_of_cycles(n_cycles, bmap);
  for (int i = 0; i < 4 * NR_BLOCKS; i++) blocks[i] = 0.0;
  int64_t count = 0;
  double volume = s->props[6];
  const int nc = m->n_sites > 1 ? m->k_or * m->n_sites * 3 : 0;
  for (int64_t c = 0; c < n_cycles; c++) {
    int b = bmap[c];
    (void)orc_mt64_uniform(&st.g); /* SelectedSystem (src/monte_carlo.c:626) */
    for (int j = 0; j < steps_per_cycle; j++) {
      (void)orc_mt64_uniform(&st.g);       /* CurrentComponent (src/monte_carlo.c:643) */
      double ran = orc_mt64_uniform(&st.g); /* move selection (src/monte_carlo.c:649) */
      if (ran < 1.0) {
        double rec[ORC_REC_N];
        orc_trials tr;
        memset(&tr, 0, sizeof(tr));
        /* WidomIdealGasAccumulated: weight * Volume / (NumberOfMolecules + 1)
           with the integer molecules present (src/mc_moves.c:7861-7863) */
        blocks[4 * b + 0] += 1.0;
        blocks[4 * b + 3] += 1.0 * (volume / ((double)s->ads_nmol + 1.0));
        orc_widom_insertion(s, m, mt_next, &st, rec, &tr);
        if (rec[ORC_REC_OVERLAP] == 0.0) {
          blocks[4 * b + 1] += 1.0 * rec[ORC_REC_W];
          blocks[4 * b + 2] += 1.0 * (0.0 + rec[ORC_REC_DU]) * rec[ORC_REC_W];
        }
        if (count < max_ins) {
          /* what the insertion drew, in the layout of host-supplied trials */
          if (recs) memcpy(recs + (size_t)count * ORC_REC_N, rec, sizeof(rec));
          if (fb)
            for (int t = 0; t < m->k_fb; t++)
              for (int d = 0; d < 3; d++) fb[((size_t)count * m->k_fb + t) * 3 + d] = tr.fb[t][d];
          if (cord && nc)
            for (int t = 0; t < m->k_or; t++)
              for (int a = 0; a < m->n_sites; a++)
                for (int d = 0; d < 3; d++)
                  cord[(size_t)count * nc + (t * m->n_sites + a) * 3 + d] = tr.cord[t][a][d];
          if (u_sel) {
            u_sel[2 * count] = tr.u_sel_fb;
            u_sel[2 * count + 1] = tr.u_sel_or;
          }
        }
        count++;
      }
    }
  }
  free(bmap);
  return count;
}

int64_t orc_widom_run_mt(const orc_system *s, const orc_molecule *m, uint64_t seed, int64_t n_cycles,
                         int steps_per_cycle, double *blocks) {
  return orc_widom_run_mt_trials(s, m, seed, n_cycles, steps_per_cycle, blocks, 0, NULL, NULL, NULL, NULL);
}

double orc_framework_density(double mass_g_mol, double volume_A3) {
  return 1e-3 * mass_g_mol / (volume_A3 * CUBE(ANGSTROM_) * AVOGADRO_CONSTANT_);
}

/* src/statistics.c:2539-2560, 2708-2722, 4747-4800 */
void orc_widom_estimators(const double *blocks, double T, double rho, double igw, double out[16]) {
  units_t u = make_units();
  double sum = 0.0, sum2 = 0.0;
  for (int i = 0; i < NR_BLOCKS; i++) {
    double tmp = 0.0;
    if (blocks[4 * i] > 0.0) {
      tmp = (1.0 / (MOLAR_GAS_CONSTANT_ * T * rho)) * blocks[4 * i + 1] / (igw * blocks[4 * i]);
      sum += tmp;
      sum2 += SQR(tmp);
    }
    out[i] = tmp;
  }
  out[5] = sum / (double)NR_BLOCKS;
  out[6] = 2.0 * sqrt(fabs((sum2 / (double)NR_BLOCKS) - SQR(sum) / (double)SQR(NR_BLOCKS)));
  sum = sum2 = 0.0;
  for (int i = 0; i < NR_BLOCKS; i++) {
    double tmp = 0.0;
    if (blocks[4 * i] > 0.0) {
      tmp = (blocks[4 * i + 2] / blocks[4 * i + 1] - 0.0 / blocks[4 * i]) * u.E2K;
      sum += tmp;
      sum2 += SQR(tmp);
    }
    out[7 + i] = tmp;
  }
  out[12] = sum / (double)NR_BLOCKS;
  out[13] = 2.0 * sqrt(fabs((sum2 / (double)NR_BLOCKS) - SQR(sum) / (double)SQR(NR_BLOCKS)));
  double c = 0.0, sw = 0.0;
  for (int i = 0; i < NR_BLOCKS; i++)
    if (blocks[4 * i] > 0.0) {
      sw += blocks[4 * i + 1] / igw;
      c += blocks[4 * i];
    }
  out[14] = c > 0.0 ? (1.0 / (MOLAR_GAS_CONSTANT_ * T * rho)) * (sw / c) : 0.0;
  out[15] = c > 0.0 ? sw * igw / c : 0.0;
}

/* ===================================================================== */
/* grids (src/grids.c:431-742, 979-1077; src/framework_hessian.c:63-109;  */
/*        src/potentials.c:10978-10994)                                   */
/* ===================================================================== */
void orc_grid_geometry(const double ub[9], double spacing, double out[12]) {
  double size[3] = {0.0, 0.0, 0.0}, shift[3] = {0.0, 0.0, 0.0};
  const int col[3][3] = {{AX, AY, AZ}, {BX, BY, BZ}, {CX, CY, CZ}};
  for (int v = 0; v < 3; v++)
    for (int d = 0; d < 3; d++) {
      size[d] += fabs(ub[col[v][d]]);
      if (ub[col[v][d]] < 0.0) shift[d] += ub[col[v][d]];
    }
  int N[3];
  for (int d = 0; d < 3; d++) {
    N[d] = (int)(size[d] / spacing);
    if (N[d] % 2 == 0) N[d]++;
  }
  for (int d = 0; d < 3; d++) {
    out[d] = (double)N[d];
    out[3 + d] = size[d];
    out[6 + d] = shift[d];
    out[9 + d] = size[d] / (double)N[d];
  }
}

void orc_derivatives_vdw_at(const orc_system *s, int type, const double pos[3], double out[14]) {
  double value = 0, fx = 0, fy = 0, fz = 0;
  double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double third = 0.0;
  const double *pp = s->pp[type];
  for (int i = 0; i < s->n; i++) {
    double dr[3] = {pos[0] - s->x[i], pos[1] - s->y[i], pos[2] - s->z[i]};
    orc_min_image(s->box, s->inv, s->bc, dr);
    double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
    if (rr < s->rc_vdw2) {
      double arg1 = pp[3 * i], arg2 = SQR(pp[3 * i + 1]), arg3 = pp[3 * i + 2];
      double rri3 = CUBE(arg2 / rr);
      double F = 4.0 * arg1 * (rri3 * (rri3 - 1.0)) - arg3;
      double DF = 24.0 * arg1 * (rri3 * (1.0 - 2.0 * rri3)) / rr;
      double DDF = 96.0 * arg1 * (rri3 * (7.0 * rri3 - 2.0)) / SQR(rr);
      double DDDF = 384.0 * arg1 * (rri3 * (5.0 - 28.0 * rri3)) / SQR(SQR(rr));
      double dx = dr[0], dy = dr[1], dz = dr[2];
      value += F;
      fx += dx * DF;
      fy += dy * DF;
      fz += dz * DF;
      h[AX] += DDF * dx * dx + DF; h[BX] += DDF * dy * dx;      h[CX] += DDF * dz * dx;
      h[AY] += DDF * dx * dy;      h[BY] += DDF * dy * dy + DF; h[CY] += DDF * dz * dy;
      h[AZ] += DDF * dx * dz;      h[BZ] += DDF * dy * dz;      h[CZ] += DDF * dz * dz + DF;
      third += DDDF * dx * dy * dz;
    }
  }
  out[0] = value;
  out[1] = fx;
  out[2] = fy;
  out[3] = fz;
  for (int i = 0; i < 9; i++) out[4 + i] = h[i];
  out[13] = third;
}

void orc_grid_node(const orc_system *s, int type, const double g[12], int i, int j, int k, float out[8]) {
  double pos[3];
  double d[14];
  pos[0] = i * g[3] / (int)g[0] + g[6];
  pos[1] = j * g[4] / (int)g[1] + g[7];
  pos[2] = k * g[5] / (int)g[2] + g[8];
  orc_derivatives_vdw_at(s, type, pos, d);
  double value = d[0];
  double f[3] = {d[1], d[2], d[3]};
  double hay = d[4 + AY], haz = d[4 + AZ], hbz = d[4 + BZ];
  double third = d[13];
  const double crit = ENERGY_OVERLAP_CRITERIA;
  if (value > crit) {
    value = 2.0 * crit;
    for (int a = 0; a < 3; a++) {
      if (f[a] > crit) f[a] = crit;
      if (f[a] < -crit) f[a] = -crit;
    }
    hay = haz = hbz = 0.0;
    third = 0.0;
  }
  double dx = g[9], dy = g[10], dz = g[11];
  out[0] = (float)value;
  out[1] = (float)((double)(float)f[0] * dx);
  out[2] = (float)((double)(float)f[1] * dy);
  out[3] = (float)((double)(float)f[2] * dz);
  out[4] = (float)((double)(float)hay * (dx * dy));
  out[5] = (float)((double)(float)haz * (dx * dz));
  out[6] = (float)((double)(float)hbz * (dy * dz));
  out[7] = (float)((double)(float)third * (dx * dy * dz));
}

/* CalculateDerivativesAtPositionReal (src/framework_hessian.c:111-175): the
 * real-space Ewald potential of a unit charge and its derivatives, in the
 * same order as out[] of orc_derivatives_vdw_at.  Returns the smallest
 * distance to an atom within the cut-off. */
double orc_derivatives_real_at(const orc_system *s, const double pos[3], double out[14]) {
  double value = 0, fx = 0, fy = 0, fz = 0;
  double h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double third = 0.0;
  const double C = s->u.C, alpha = s->alpha, ChargeA = 1.0;
  double smallest_r = DBL_MAX;
  for (int i = 0; i < s->n; i++) {
    double ChargeB = s->q[i];
    double dr[3] = {pos[0] - s->x[i], pos[1] - s->y[i], pos[2] - s->z[i]};
    orc_min_image(s->box, s->inv, s->bc, dr);
    double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
    if (rr < s->rc_charge2) {
      double r = sqrt(rr);
      if (r < smallest_r) smallest_r = r;
      double F = C * ChargeA * ChargeB * (erfc(alpha * r) / r);
      double DF = -C * ChargeA * ChargeB *
                  (2.0 * alpha * r * exp(-SQR(alpha) * rr) / sqrt(M_PI) + erfc(alpha * r)) / (r * rr);
      double DDF = ChargeA * ChargeB * C *
                   (2.0 * alpha * exp(-SQR(alpha) * rr) * r * (3.0 + 2.0 * SQR(alpha) * rr) / sqrt(M_PI) +
                    3.0 * erfc(alpha * r)) /
                   (rr * rr * r);
      double DDDF = ChargeA * ChargeB * C *
                    (-2.0 * alpha * exp(-SQR(alpha) * rr) * r *
                         (15.0 + 10.0 * SQR(alpha) * rr + 4.0 * SQR(SQR(alpha) * rr)) / sqrt(M_PI) -
                     15.0 * erfc(alpha * r)) /
                    (rr * rr * rr * r);
      double dx = dr[0], dy = dr[1], dz = dr[2];
      value += F;
      fx += dx * DF;
      fy += dy * DF;
      fz += dz * DF;
      h[AX] += DDF * dx * dx + DF; h[BX] += DDF * dy * dx;      h[CX] += DDF * dz * dx;
      h[AY] += DDF * dx * dy;      h[BY] += DDF * dy * dy + DF; h[CY] += DDF * dz * dy;
      h[AZ] += DDF * dx * dz;      h[BZ] += DDF * dy * dz;      h[CZ] += DDF * dz * dz + DF;
      third += DDDF * dx * dy * dz;
    }
  }
  out[0] = value;
  out[1] = fx;
  out[2] = fy;
  out[3] = fz;
  for (int i = 0; i < 9; i++) out[4 + i] = h[i];
  out[13] = third;
  return smallest_r;
}

/* A node of the Ewald-Coulomb grid (src/grids.c:634-683): capped when the
 * energy exceeds the overlap criterion or an atom is closer than 1 A. */
void orc_coulomb_grid_node(const orc_system *s, const double g[12], int i, int j, int k, float out[8]) {
  double pos[3];
  double d[14];
  pos[0] = i * g[3] / (int)g[0] + g[6];
  pos[1] = j * g[4] / (int)g[1] + g[7];
  pos[2] = k * g[5] / (int)g[2] + g[8];
  double smallest_r = orc_derivatives_real_at(s, pos, d);
  double value = d[0];
  double f[3] = {d[1], d[2], d[3]};
  double hay = d[4 + AY], haz = d[4 + AZ], hbz = d[4 + BZ];
  double third = d[13];
  const double crit = ENERGY_OVERLAP_CRITERIA;
  if (value > crit || smallest_r < 1.0) {
    value = 2.0 * crit;
    for (int a = 0; a < 3; a++) {
      if (f[a] > crit) f[a] = crit;
      if (f[a] < -crit) f[a] = -crit;
    }
    hay = haz = hbz = 0.0;
    third = 0.0;
  }
  double dx = g[9], dy = g[10], dz = g[11];
  out[0] = (float)value;
  out[1] = (float)((double)(float)f[0] * dx);
  out[2] = (float)((double)(float)f[1] * dy);
  out[3] = (float)((double)(float)f[2] * dz);
  out[4] = (float)((double)(float)hay * (dx * dy));
  out[5] = (float)((double)(float)haz * (dx * dz));
  out[6] = (float)((double)(float)hbz * (dy * dz));
  out[7] = (float)((double)(float)third * (dx * dy * dz));
}

/* Lekien & Marsden (2005): the 64x64 matrix mapping corner values and
 * derivatives (f, fx, fy, fz, fxy, fxz, fyz, fxyz at the 8 corners, row
 * corner+8*derivative) to the coefficients a[i+4j+16k] of x^i y^j z^k.
 * Generated here as the inverse of the constraint matrix. */
void orc_lm_coeff(int out[64 * 64]) {
  static const int dd[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                               {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  double B[64][128];
  for (int row = 0; row < 64; row++) {
    int c = row % 8, d = row / 8;
    int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    for (int col = 0; col < 64; col++) {
      int p[3] = {col % 4, (col / 4) % 4, col / 16};
      int cc[3] = {cx, cy, cz};
      double v = 1.0;
      for (int a = 0; a < 3; a++) {
        int pw = p[a], der = dd[d][a], at = cc[a];
        double t;
        if (der == 0)
          t = (pw == 0) ? 1.0 : (at ? 1.0 : 0.0);
        else
          t = (pw == 0) ? 0.0 : (pw == 1 ? 1.0 : (at ? (double)pw : 0.0));
        v *= t;
      }
      B[row][col] = v;
      B[row][64 + col] = (row == col) ? 1.0 : 0.0;
    }
  }
  for (int col = 0; col < 64; col++) {
    int piv = col;
    for (int r = col + 1; r < 64; r++)
      if (fabs(B[r][col]) > fabs(B[piv][col])) piv = r;
    if (piv != col)
      for (int c = 0; c < 128; c++) {
        double t = B[col][c];
        B[col][c] = B[piv][c];
        B[piv][c] = t;
      }
    double p = B[col][col];
    for (int c = 0; c < 128; c++) B[col][c] /= p;
    for (int r = 0; r < 64; r++)
      if (r != col && B[r][col] != 0.0) {
        double f = B[r][col];
        for (int c = 0; c < 128; c++) B[r][c] -= f * B[col][c];
      }
  }
  for (int r = 0; r < 64; r++)
    for (int c = 0; c < 64; c++) out[r * 64 + c] = (int)lround(B[r][64 + c]);
}

/* Position -> grid cell and fractional offsets (src/grids.c:985-1012). */
static void grid_locate(const double g[12], int bc, const double ub[9], const double uinv[9],
                        const double pos_in[3], int x0[3], double r[3]) {
  double pos[3] = {pos_in[0], pos_in[1], pos_in[2]};
  double sv[3];
  int N[3] = {(int)g[0], (int)g[1], (int)g[2]};
  if (bc == ORC_BC_RECTANGULAR) {
    /* UnitCellSize = lengths of the unit-cell vectors */
    double L[3] = {ub[AX], ub[BY], ub[CZ]};
    for (int d = 0; d < 3; d++) {
      pos[d] -= L[d] * (double)(nint(pos[d] / L[d]));
      if (pos[d] < 0.0) pos[d] += L[d];
    }
  } else {
    sv[0] = uinv[AX] * pos[0] + uinv[BX] * pos[1] + uinv[CX] * pos[2];
    sv[1] = uinv[AY] * pos[0] + uinv[BY] * pos[1] + uinv[CY] * pos[2];
    sv[2] = uinv[AZ] * pos[0] + uinv[BZ] * pos[1] + uinv[CZ] * pos[2];
    for (int d = 0; d < 3; d++) {
      sv[d] -= (double)nint(sv[d]);
      if (sv[d] < 0.0) sv[d] += 1.0;
    }
    pos[0] = ub[AX] * sv[0] + ub[BX] * sv[1] + ub[CX] * sv[2];
    pos[1] = ub[AY] * sv[0] + ub[BY] * sv[1] + ub[CY] * sv[2];
    pos[2] = ub[AZ] * sv[0] + ub[BZ] * sv[1] + ub[CZ] * sv[2];
  }
  double s3[3];
  for (int d = 0; d < 3; d++) s3[d] = ((pos[d] - g[6 + d]) / g[3 + d]) * (double)(N[d]);
  for (int d = 0; d < 3; d++) {
    x0[d] = (int)s3[d];
    r[d] = s3[d] - x0[d];
  }
}

/* Tricubic value from the 8 channels of the 8 corners, X[corner + 8 *
 * channel], with the Lekien-Marsden matrix (src/grids.c:1014-1077); cap
 * selects the overlap short-cut of the VDW grid (src/grids.c:1058-1060),
 * which InterpolateCoulombGrid does not have. */
static double grid_tricubic_cap(const double X[64], const double r[3], int cap) {
  static int Coeff[64 * 64];
  static int have = 0;
  if (!have) {
    orc_lm_coeff(Coeff);
    have = 1;
  }
  double a[64];
  if (cap)
    for (int i = 0; i < 8; i++)
      if (X[i] > 1e10) return 1e8;
  for (int i = 0; i < 64; i++) {
    a[i] = (double)(0.0);
    for (int j = 0; j < 64; j++) a[i] += Coeff[i * 64 + j] * X[j];
  }
  double value = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++) value += a[i + 4 * j + 16 * k] * pow(r[0], i) * pow(r[1], j) * pow(r[2], k);
  return value;
}

static double grid_tricubic(const double X[64], const double r[3]) { return grid_tricubic_cap(X, r, 1); }

/* InterpolateCoulombGrid (src/grids.c:1200-1295): the same wrap and
 * tricubic without the overlap short-cut, times the probe charge. */
double orc_interpolate_coulomb_grid(const float *grid, const double g[12], int bc, const double ub[9],
                                    const double uinv[9], const double pos_in[3], double charge) {
  int x[3];
  double r[3];
  grid_locate(g, bc, ub, uinv, pos_in, x, r);
  int ny = (int)g[1] + 1, nz = (int)g[2] + 1;
  double X[64];
  for (int c = 0; c < 8; c++) {
    const float *node = grid + 8 * (((size_t)(x[0] + (c & 1)) * ny + (x[1] + ((c >> 1) & 1))) * nz +
                                    (x[2] + ((c >> 2) & 1)));
    for (int i = 0; i < 8; i++) X[c + i * 8] = node[i];
  }
  return charge * grid_tricubic_cap(X, r, 0);
}

double orc_interpolate_grid(const float *grid, const double g[12], int bc, const double ub[9],
                            const double uinv[9], const double pos_in[3]) {
  int x[3];
  double r[3];
  grid_locate(g, bc, ub, uinv, pos_in, x, r);
  int ny = (int)g[1] + 1, nz = (int)g[2] + 1;
  double X[64];
  for (int c = 0; c < 8; c++) {
    const float *node = grid + 8 * (((size_t)(x[0] + (c & 1)) * ny + (x[1] + ((c >> 1) & 1))) * nz +
                                    (x[2] + ((c >> 2) & 1)));
    for (int i = 0; i < 8; i++) X[c + i * 8] = node[i];
  }
  return grid_tricubic(X, r);
}

double orc_interpolate_direct(const orc_system *s, int type, const double g[12], int bc,
                              const double ub[9], const double uinv[9], const double pos_in[3]) {
  int x[3];
  double r[3];
  grid_locate(g, bc, ub, uinv, pos_in, x, r);
  double X[64];
  for (int c = 0; c < 8; c++) {
    float node[8];
    orc_grid_node(s, type, g, x[0] + (c & 1), x[1] + ((c >> 1) & 1), x[2] + ((c >> 2) & 1), node);
    for (int i = 0; i < 8; i++) X[c + i * 8] = node[i];
  }
  return grid_tricubic(X, r);
}

/* ===================================================================== */
/* Flexible chains: CBMC growth with the internal MC scheme               */
/* (src/cbmc.c:623-795, 1154-2482, 2582-2849, 3056-3185,                  */
/*  src/internal_energy.c:48-490, 1168-1340, 2177-2430,                   */
/*  src/utils.c:640-990, src/mc_moves.c:7844-7946)                        */
/* ===================================================================== */
#define CH_FIXED_BOND 0
#define CH_HARMONIC_BOND 1
#define CH_MAX_TORSION_TRIALS 1024
#define DEG2RAD_ (M_PI / 180.0) /* src/constants.h:108 */
#define SIGN_(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a)) /* src/utils.h:65 */

typedef struct {
  int cur, prev, n_prev;
  int n_place, place[ORC_CHAIN_MAX];
  int n_branch, branch[ORC_CHAIN_MAX];
  int n_bonds, bonds[ORC_CHAIN_MAX];
  int n_bends, bends[2 * ORC_CHAIN_MAX];
  int n_tors, tors[4 * ORC_CHAIN_MAX];
  int n_intra, intra[ORC_CHAIN_MAX_INTRA];
} chain_step;

/* SetConnectivityMatrix, SetGrowingStatus and Interactions of a molecule
 * made of one flexible group (src/cbmc.c:2582-2760, 901-1152); the first
 * bead (StartingBead) is placed.  The growth path is fixed (the first
 * possible bead, src/cbmc.c:2651), so the steps are the same for every
 * insertion. */
static int chain_steps(const orc_chain *c, chain_step *steps) {
  const int n = c->n_atoms;
  static char conn[ORC_CHAIN_MAX][ORC_CHAIN_MAX], todo[ORC_CHAIN_MAX][ORC_CHAIN_MAX];
  memset(conn, 0, sizeof(conn));
  for (int j = 0; j < c->n_bonds; j++) {
    conn[c->bond[j][0]][c->bond[j][1]] = 1;
    conn[c->bond[j][1]][c->bond[j][0]] = 1;
  }
  memcpy(todo, conn, sizeof(conn));
  int placed[ORC_CHAIN_MAX], n_placed = 1;
  placed[0] = c->start;
  for (int j = 0; j < n; j++) todo[c->start][j] = 0;
  int ns = 0;
  while (n_placed != n) {
    chain_step *st = &steps[ns];
    memset(st, 0, sizeof(*st));
    /* SetGrowingStatus: the first (placed bead, unplaced neighbour) pair */
    int beadn = -1, curb = -1;
    for (int i = 0; i < n_placed && beadn < 0; i++) {
      int k = placed[i];
      for (int j = 0; j < n; j++)
        if (todo[j][k]) {
          beadn = j;
          curb = k;
          break;
        }
    }
    if (beadn < 0) return -1; /* not connected */
    st->cur = curb;
    st->prev = -1;
    int attached = 0;
    for (int i = 0; i < n; i++)
      if (conn[i][curb]) {
        if (todo[i][curb])
          attached = 1;
        else {
          st->n_prev++;
          st->prev = i;
        }
      }
    /* the flexible group adds every unplaced atom bonded to the current
     * bead, each as a branch of its own */
    if (attached)
      for (int j = 0; j < n; j++)
        if (todo[j][curb]) {
          st->place[st->n_place++] = j;
          st->branch[st->n_branch++] = j;
        }
    if (st->n_prev == 0) {
      st->n_place = 1;
      st->place[0] = beadn;
    }
    if (st->n_prev > 1) return -1;
    /* Interactions */
    char done[ORC_CHAIN_MAX] = {0}, tbp[ORC_CHAIN_MAX] = {0};
    for (int i = 0; i < n_placed; i++) done[placed[i]] = 1;
    for (int i = 0; i < st->n_place; i++) {
      tbp[st->place[i]] = 1;
      done[st->place[i]] = 1;
    }
    for (int i = 0; i < st->n_place; i++) {
      int A = st->place[i] < curb ? st->place[i] : curb, B = st->place[i] < curb ? curb : st->place[i];
      for (int j = 0; j < c->n_bonds; j++)
        if ((A == c->bond[j][0] && B == c->bond[j][1]) || (A == c->bond[j][1] && B == c->bond[j][0]))
          st->bonds[st->n_bonds++] = j;
    }
    for (int j = 0; j < c->n_bends; j++) {
      const int *b = c->bend[j];
      if (done[b[0]] && done[b[1]] && done[b[2]] && (tbp[b[0]] || tbp[b[1]] || tbp[b[2]]))
        st->bends[st->n_bends++] = j;
    }
    for (int j = 0; j < c->n_torsions; j++) {
      const int *t = c->torsion[j];
      if (done[t[0]] && done[t[1]] && done[t[2]] && done[t[3]] && (tbp[t[0]] || tbp[t[1]] || tbp[t[2]] || tbp[t[3]]))
        st->tors[st->n_tors++] = j;
    }
    for (int j = 0; j < c->n_intra; j++) {
      const int *p = c->intra[j];
      if (done[p[0]] && done[p[1]] && (tbp[p[0]] || tbp[p[1]])) st->intra[st->n_intra++] = j;
    }
    if (st->n_prev == 0 && st->n_bonds == 0) return -1; /* nothing to grow the bead along */
    for (int i = 0; i < st->n_place; i++) {
      int j = st->place[i];
      placed[n_placed++] = j;
      for (int ip = 0; ip < n; ip++) todo[j][ip] = 0;
    }
    ns++;
  }
  return ns;
}

int orc_chain_plan(const orc_chain *c, int *info, int *place) {
  chain_step *steps = (chain_step *)malloc(sizeof(chain_step) * ORC_CHAIN_MAX);
  int ns = chain_steps(c, steps);
  for (int i = 0; i < ns; i++) {
    const chain_step *st = &steps[i];
    int *o = info + 8 * i;
    o[0] = st->cur;
    o[1] = st->prev;
    o[2] = st->n_prev;
    o[3] = st->n_place;
    o[4] = st->n_bonds;
    o[5] = st->n_bends;
    o[6] = st->n_tors;
    o[7] = st->n_intra;
    for (int j = 0; j < st->n_place; j++) place[ORC_CHAIN_MAX * i + j] = st->place[j];
  }
  free(steps);
  return ns;
}

/* parameters as the reference holds them after reading the molecule file
 * (src/molecule.c:1638-1644, 1790-1796, 2173-2184): force constants divided
 * by ENERGY_TO_KELVIN, angles in radians; LJ of the intra pairs as
 * PotentialValue mixes them, and the 1-4 scaling (src/molecule.c:2997-3025) */
typedef struct {
  double bond[ORC_CHAIN_MAX][4];
  double bend[2 * ORC_CHAIN_MAX][4];
  double tors[4 * ORC_CHAIN_MAX][4];
  double scal[ORC_CHAIN_MAX_INTRA];
  double lj[ORC_CHAIN_MAX_INTRA][3];
} chain_par;

static void chain_params(const orc_system *s, const orc_chain *c, chain_par *p) {
  for (int j = 0; j < c->n_bonds; j++) {
    for (int q = 0; q < 4; q++) p->bond[j][q] = c->bond_par[j][q];
    if (c->bond[j][2] == CH_HARMONIC_BOND) p->bond[j][0] /= s->u.E2K;
  }
  for (int j = 0; j < c->n_bends; j++) {
    for (int q = 0; q < 4; q++) p->bend[j][q] = c->bend_par[j][q];
    p->bend[j][0] /= s->u.E2K;
    p->bend[j][1] *= DEG2RAD_;
  }
  for (int j = 0; j < c->n_torsions; j++)
    for (int q = 0; q < 4; q++) p->tors[j][q] = c->torsion_par[j][q] / s->u.E2K;
  for (int i = 0; i < c->n_intra; i++) {
    int A = c->intra[i][0], B = c->intra[i][1];
    p->scal[i] = 1.0;
    for (int j = 0; j < c->n_bends; j++) {
      int A1 = c->bend[j][0], B1 = c->bend[j][1], C1 = c->bend[j][2];
      for (int k = 0; k < c->n_bends; k++) {
        int A2 = c->bend[k][0], B2 = c->bend[k][1], C2 = c->bend[k][2];
        if ((B1 == A2) && (C1 == B2) && (((A1 == A) && (C2 == B)) || ((A1 == B) && (C2 == A)))) p->scal[i] = c->intra14_scaling;
        if ((B1 == C2) && (C1 == B2) && (((A1 == A) && (A2 == B)) || ((A1 == B) && (A2 == A)))) p->scal[i] = c->intra14_scaling;
        if ((A1 == B2) && (B1 == A2) && (((C1 == A) && (C2 == B)) || ((C1 == B) && (C2 == A)))) p->scal[i] = c->intra14_scaling;
        if ((A1 == B2) && (B1 == C2) && (((C1 == A) && (A2 == B)) || ((C1 == B) && (A2 == A)))) p->scal[i] = c->intra14_scaling;
      }
    }
    pair_lj(s, c->type[A], c->type[B], p->lj[i]);
  }
}

/* CalculateBondEnergy (src/internal_energy.c:372-490), HARMONIC_BOND and FIXED_BOND */
static double ch_bond_u(int kind, const double *par, const double a[3], const double b[3]) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  double rr = SQR(dx) + SQR(dy) + SQR(dz);
  double r = sqrt(rr);
  if (kind == CH_HARMONIC_BOND) return 0.5 * par[0] * SQR(r - par[1]);
  return 0.0;
}

/* CalculateBendEnergy (src/internal_energy.c:1168-1340), HARMONIC_BEND */
static double ch_bend_u(const double *par, const double a[3], const double b[3], const double c[3]) {
  double Rab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  double rab = sqrt(SQR(Rab[0]) + SQR(Rab[1]) + SQR(Rab[2]));
  Rab[0] /= rab;
  Rab[1] /= rab;
  Rab[2] /= rab;
  double Rbc[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  double rbc = sqrt(SQR(Rbc[0]) + SQR(Rbc[1]) + SQR(Rbc[2]));
  Rbc[0] /= rbc;
  Rbc[1] /= rbc;
  Rbc[2] /= rbc;
  double CosTheta = (Rab[0] * Rbc[0] + Rab[1] * Rbc[1] + Rab[2] * Rbc[2]);
  const double lo = (-1.0 > CosTheta) ? -1.0 : CosTheta; /* MIN2(1.0,MAX2(-1.0,CosTheta)), src/utils.h:63-64 */
  CosTheta = (1.0 <= lo) ? 1.0 : lo;
  double Theta = acos(CosTheta);
  return 0.5 * par[0] * SQR(Theta - par[1]);
}

/* CalculateTorsionEnergy (src/internal_energy.c:2177-2330), TRAPPE_DIHEDRAL */
static double ch_torsion_u(const double *par, const double a[3], const double b[3], const double c[3],
                           const double d[3]) {
  double Dab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  double Dbc[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  double rbc = sqrt(SQR(Dbc[0]) + SQR(Dbc[1]) + SQR(Dbc[2]));
  Dbc[0] /= rbc;
  Dbc[1] /= rbc;
  Dbc[2] /= rbc;
  double Dcd[3] = {d[0] - c[0], d[1] - c[1], d[2] - c[2]};
  double dot_ab = Dab[0] * Dbc[0] + Dab[1] * Dbc[1] + Dab[2] * Dbc[2];
  double dot_cd = Dcd[0] * Dbc[0] + Dcd[1] * Dbc[1] + Dcd[2] * Dbc[2];
  double dr[3] = {Dab[0] - dot_ab * Dbc[0], Dab[1] - dot_ab * Dbc[1], Dab[2] - dot_ab * Dbc[2]};
  double r = sqrt(SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]));
  dr[0] /= r;
  dr[1] /= r;
  dr[2] /= r;
  double ds[3] = {Dcd[0] - dot_cd * Dbc[0], Dcd[1] - dot_cd * Dbc[1], Dcd[2] - dot_cd * Dbc[2]};
  double sv = sqrt(SQR(ds[0]) + SQR(ds[1]) + SQR(ds[2]));
  ds[0] /= sv;
  ds[1] /= sv;
  ds[2] /= sv;
  double CosPhi = dr[0] * ds[0] + dr[1] * ds[1] + dr[2] * ds[2];
  double m = fabs(CosPhi) <= 1.0 ? fabs(CosPhi) : 1.0;
  CosPhi = SIGN_(m, CosPhi);
  return par[0] + (1.0 + CosPhi) * (par[1] + par[3] - 2.0 * (CosPhi - 1.0) * (par[2] - 2.0 * par[3] * CosPhi));
}

double orc_chain_bond_energy(const orc_system *s, int kind, const double par[4], double r) {
  double p[4] = {par[0], par[1], par[2], par[3]};
  if (kind == CH_HARMONIC_BOND) p[0] /= s->u.E2K;
  double a[3] = {0.0, 0.0, 0.0}, b[3] = {r, 0.0, 0.0};
  return ch_bond_u(kind, p, a, b);
}

double orc_chain_bend_energy(const orc_system *s, int kind, const double par[4], const double a[3],
                             const double b[3], const double c[3]) {
  (void)kind;
  double p[4] = {par[0] / s->u.E2K, par[1] * DEG2RAD_, par[2], par[3]};
  return ch_bend_u(p, a, b, c);
}

double orc_chain_torsion_energy(const orc_system *s, int kind, const double par[4], const double a[3],
                                const double b[3], const double c[3], const double d[3]) {
  (void)kind;
  double p[4] = {par[0] / s->u.E2K, par[1] / s->u.E2K, par[2] / s->u.E2K, par[3] / s->u.E2K};
  return ch_torsion_u(p, a, b, c, d);
}

/* draws: one MT19937-64 stream in the reference's order, or the counter
 * stream with one sub-stream per purpose */
typedef struct {
  orc_mt64 *g;   /* non-null: MT mode */
  uint64_t key;  /* insertion key */
  uint64_t sub;  /* current sub-stream key */
  uint64_t ctr;
} cdraw_t;

uint64_t orc_chain_substream(uint64_t insertion_key, uint64_t tag) {
  return mix64(insertion_key ^ mix64(tag + 0x2545F4914F6CDD1DULL));
}
static void cd_sub(cdraw_t *d, uint64_t tag) {
  if (d->g) return;
  d->sub = orc_chain_substream(d->key, tag);
  d->ctr = 0;
}
static double cd(cdraw_t *d) {
  return d->g ? orc_mt64_uniform(d->g) : orc_stream_uniform(d->sub, d->ctr++);
}
static uint64_t ch_tag(int step, int j) { return ((uint64_t)step << 32) | (uint64_t)j; }

/* RandomGaussianNumber (src/utils.c:640-651) */
static double ch_gauss(cdraw_t *d) {
  double ran1, ran2, r2;
  do {
    ran1 = 2.0 * cd(d) - 1.0;
    ran2 = 2.0 * cd(d) - 1.0;
    r2 = SQR(ran1) + SQR(ran2);
  } while ((r2 > 1.0) || (r2 == 0.0));
  return ran2 * sqrt(-2.0 * log(r2) / r2);
}

/* RandomNumberOnUnitSphere (src/utils.c:688-708) */
static void ch_sphere(cdraw_t *d, double v[3]) {
  double ran1, ran2, ransq;
  do {
    ran1 = 2.0 * cd(d) - 1.0;
    ran2 = 2.0 * cd(d) - 1.0;
    ransq = SQR(ran1) + SQR(ran2);
  } while (ransq >= 1.0);
  double ranh = 2.0 * sqrt(1.0 - ransq);
  v[0] = ran1 * ranh;
  v[1] = ran2 * ranh;
  v[2] = 1.0 - 2.0 * ransq;
}

/* RandomNumberOnCone (src/utils.c:709-735) */
static void ch_cone(cdraw_t *d, const double vin[3], double theta, double v[3]) {
  double cosin = cos(theta);
  double ran1 = 1.0 / sqrt(SQR(vin[0]) + SQR(vin[1]) + SQR(vin[2]));
  double xprev = vin[0] * ran1, yprev = vin[1] * ran1, zprev = vin[2] * ran1;
  ch_sphere(d, v);
  ran1 = v[0] * xprev + v[1] * yprev + v[2] * zprev;
  v[0] = v[0] - ran1 * xprev;
  v[1] = v[1] - ran1 * yprev;
  v[2] = v[2] - ran1 * zprev;
  ran1 = sqrt((1.0 - cosin * cosin) / (SQR(v[0]) + SQR(v[1]) + SQR(v[2])));
  v[0] = v[0] * ran1 + cosin * xprev;
  v[1] = v[1] * ran1 + cosin * yprev;
  v[2] = v[2] * ran1 + cosin * zprev;
  ran1 = 1.0 / sqrt(SQR(v[0]) + SQR(v[1]) + SQR(v[2]));
  v[0] = v[0] * ran1;
  v[1] = v[1] * ran1;
  v[2] = v[2] * ran1;
}

/* Perpendicular (src/utils.c:789-832) */
static void ch_perpendicular(const double a[3], const double b[3], double v[3]) {
  double fx = fabs(b[1] * a[2] - a[1] * b[2]);
  double fy = fabs(b[0] * a[2] - a[0] * b[2]);
  double fz = fabs(b[1] * a[0] - a[1] * b[0]);
  int choise = 1;
  if (fy > fx) {
    choise = 2;
    fx = fy;
  }
  if (fz > fx) choise = 3;
  if (choise == 1) {
    v[0] = 1.0;
    v[1] = (a[0] * b[2] - b[0] * a[2]) / (b[1] * a[2] - a[1] * b[2]);
    v[2] = (b[0] * a[1] - a[0] * b[1]) / (b[1] * a[2] - a[1] * b[2]);
  } else if (choise == 2) {
    v[0] = (a[1] * b[2] - b[1] * a[2]) / (b[0] * a[2] - a[0] * b[2]);
    v[1] = 1.0;
    v[2] = (b[1] * a[0] - a[1] * b[0]) / (b[0] * a[2] - a[0] * b[2]);
  } else {
    v[0] = (b[2] * a[1] - a[2] * b[1]) / (b[1] * a[0] - a[1] * b[0]);
    v[1] = (a[2] * b[0] - b[2] * a[0]) / (b[1] * a[0] - a[1] * b[0]);
    v[2] = 1.0;
  }
  double f = 1.0 / sqrt(SQR(v[0]) + SQR(v[1]) + SQR(v[2]));
  v[0] *= f;
  v[1] *= f;
  v[2] *= f;
}

/* RotationAroundXYZAxis (src/utils.c:945-975) */
static void ch_rotate(const double v[3], double (*cord)[3], int n, double theta) {
  double c = cos(theta);
  double w = 1.0 - c;
  double s = sqrt(1.0 - c * c);
  if (theta < 0.0) s = -s;
  double r[3][3];
  r[0][0] = (v[0]) * (v[0]) * w + c;
  r[0][1] = (v[0]) * (v[1]) * w + (v[2]) * s;
  r[0][2] = (v[0]) * (v[2]) * w - (v[1]) * s;
  r[1][0] = (v[0]) * (v[1]) * w - (v[2]) * s;
  r[1][1] = (v[1]) * (v[1]) * w + c;
  r[1][2] = (v[1]) * (v[2]) * w + (v[0]) * s;
  r[2][0] = (v[0]) * (v[2]) * w + (v[1]) * s;
  r[2][1] = (v[1]) * (v[2]) * w - (v[0]) * s;
  r[2][2] = (v[2]) * (v[2]) * w + c;
  for (int i = 0; i < n; i++) {
    double x = cord[i][0] * r[0][0] + cord[i][1] * r[0][1] + cord[i][2] * r[0][2];
    double y = cord[i][0] * r[1][0] + cord[i][1] * r[1][1] + cord[i][2] * r[1][2];
    double z = cord[i][0] * r[2][0] + cord[i][1] * r[2][1] + cord[i][2] * r[2][2];
    cord[i][0] = x;
    cord[i][1] = y;
    cord[i][2] = z;
  }
}

/* CalculateAngle (src/internal_energy.c:306-329) */
static double ch_angle(const double a[3], const double b[3], const double c[3]) {
  double Rab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  double rab = 1.0 / sqrt(SQR(Rab[0]) + SQR(Rab[1]) + SQR(Rab[2]));
  Rab[0] *= rab;
  Rab[1] *= rab;
  Rab[2] *= rab;
  double Rbc[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  double rbc = 1.0 / sqrt(SQR(Rbc[0]) + SQR(Rbc[1]) + SQR(Rbc[2]));
  Rbc[0] *= rbc;
  Rbc[1] *= rbc;
  Rbc[2] *= rbc;
  return acos(Rab[0] * Rbc[0] + Rab[1] * Rbc[1] + Rab[2] * Rbc[2]);
}

/* GenerateBondlength (src/internal_energy.c:48-70, 185-187) */
static double ch_bond_length(cdraw_t *d, double beta, int kind, const double *par) {
  if (kind != CH_HARMONIC_BOND) return par[0];
  double ran1 = 1.0 / sqrt(beta * par[0]);
  double ran2 = 1.0 / SQR(par[1] + 3.0 * ran1);
  double bl;
  do
    bl = par[1] + ran1 * ch_gauss(d);
  while ((cd(d) > (SQR(bl) * ran2)) || (bl <= 0.0));
  return bl;
}

/* GetReferenceBondLength (src/cbmc.c:809-895) */
static double ch_ref_bond_length(const orc_chain *c, int A, int B) {
  for (int i = 0; i < c->n_bonds; i++)
    if ((c->bond[i][0] == A && c->bond[i][1] == B) || (c->bond[i][1] == A && c->bond[i][0] == B))
      return c->bond[i][2] == CH_HARMONIC_BOND ? c->bond_par[i][1] : c->bond_par[i][0];
  return 1.5;
}

/* SelectTrialPosition (src/cbmc.c:563-595) over up to CH_MAX_TORSION_TRIALS */
static int ch_select(const double *bf, const int *overlap, int n, double u) {
  static double shifted[CH_MAX_TORSION_TRIALS];
  double largest = -DBL_MAX, sum = 0.0;
  for (int i = 0; i < n; i++)
    if (!overlap[i])
      if (bf[i] > largest) largest = bf[i];
  for (int i = 0; i < n; i++) {
    if (!overlap[i]) {
      shifted[i] = exp(bf[i] - largest);
      sum += shifted[i];
    } else
      shifted[i] = 0.0;
  }
  int selected = 0;
  double cumw = shifted[0];
  double ws = u * sum;
  while (cumw < ws) cumw += shifted[++selected];
  return selected;
}

typedef struct {
  chain_step steps[ORC_CHAIN_MAX];
  int n_steps;
  chain_par par;
  double tpos[64][ORC_CHAIN_MAX][3]; /* TrialPositions[iu][atom] */
} chain_work;

static double ch_sum_bonds(const orc_chain *c, const chain_work *w, const chain_step *st, int iu) {
  double u = 0.0;
  for (int j = 0; j < st->n_bonds; j++) {
    int b = st->bonds[j];
    u += ch_bond_u(c->bond[b][2], w->par.bond[b], w->tpos[iu][c->bond[b][0]], w->tpos[iu][c->bond[b][1]]);
  }
  return u;
}
static double ch_sum_bends(const orc_chain *c, const chain_work *w, const chain_step *st, int iu) {
  double u = 0.0;
  for (int j = 0; j < st->n_bends; j++) {
    int b = st->bends[j];
    u += ch_bend_u(w->par.bend[b], w->tpos[iu][c->bend[b][0]], w->tpos[iu][c->bend[b][1]], w->tpos[iu][c->bend[b][2]]);
  }
  return u;
}
static double ch_sum_torsions(const orc_chain *c, const chain_work *w, const chain_step *st, int iu) {
  double u = 0.0;
  for (int j = 0; j < st->n_tors; j++) {
    int t = st->tors[j];
    u += ch_torsion_u(w->par.tors[t], w->tpos[iu][c->torsion[t][0]], w->tpos[iu][c->torsion[t][1]],
                      w->tpos[iu][c->torsion[t][2]], w->tpos[iu][c->torsion[t][3]]);
  }
  return u;
}

/* GenerateTrialOrientationsMCScheme(FALSE) without a stored configuration
 * (src/cbmc.c:1580-2482): bond lengths drawn (and unused), branches placed on
 * the 120 degree cone at their reference bond lengths, then the internal
 * Metropolis walk over bond lengths, bend angles and cone rotations of the
 * open beads on trial 0, then per trial the torsion selection among
 * k_torsion rotations about the previous-current axis.  RosenbluthTorsion
 * averages the first NumberOfTrialPositions of the torsion trials, as the
 * reference's ComputeNormalizedRosenbluthWeight reads the global count
 * (src/cbmc.c:542-555). */
static void ch_mc_scheme(const orc_system *s, const orc_chain *c, chain_work *w, const chain_step *st, int step,
                         cdraw_t *d, double *RT) {
  const double beta = s->beta;
  const int nb = st->n_bonds, nbe = st->n_bends, k = c->k_trial, kt = c->k_torsion;
  double (*t0)[3] = w->tpos[0];
  const int cur = st->cur, prev = st->prev;
  cd_sub(d, ch_tag(step, 0));
  for (int j = 0; j < nb; j++) (void)ch_bond_length(d, beta, c->bond[st->bonds[j]][2], w->par.bond[st->bonds[j]]);
  double pv[3] = {t0[prev][0] - t0[cur][0], t0[prev][1] - t0[cur][1], t0[prev][2] - t0[cur][2]};
  double normalize = sqrt(SQR(pv[0]) + SQR(pv[1]) + SQR(pv[2]));
  pv[0] /= normalize;
  pv[1] /= normalize;
  pv[2] /= normalize;
  for (int jj = 0; jj < st->n_branch; jj++) {
    double vec[3];
    ch_cone(d, pv, 120 * DEG2RAD_, vec);
    double bl = ch_ref_bond_length(c, cur, st->branch[jj]);
    int a = st->branch[jj];
    t0[a][0] = t0[cur][0] + bl * vec[0];
    t0[a][1] = t0[cur][1] + bl * vec[1];
    t0[a][2] = t0[cur][2] + bl * vec[2];
  }
  double eobond[ORC_CHAIN_MAX], enbond[ORC_CHAIN_MAX], eobend[2 * ORC_CHAIN_MAX], enbend[2 * ORC_CHAIN_MAX];
#define CH_BOND_E(j) ch_bond_u(c->bond[st->bonds[j]][2], w->par.bond[st->bonds[j]], t0[c->bond[st->bonds[j]][0]], t0[c->bond[st->bonds[j]][1]])
#define CH_BEND_E(j) ch_bend_u(w->par.bend[st->bends[j]], t0[c->bend[st->bends[j]][0]], t0[c->bend[st->bends[j]][1]], t0[c->bend[st->bends[j]][2]])
  for (int j = 0; j < nb; j++) eobond[j] = enbond[j] = CH_BOND_E(j);
  for (int j = 0; j < nbe; j++) eobend[j] = enbend[j] = CH_BEND_E(j);
  const int trialmove = 4 * c->moves * st->n_branch;
  /* no chiral centres: swap two branches with probability 1/2 */
  if ((nb >= 2) && (cd(d) < 0.5)) {
    int A = st->branch[0], C = st->branch[1];
    double cA[3], cC[3];
    for (int q = 0; q < 3; q++) {
      cA[q] = t0[A][q] - t0[cur][q];
      cC[q] = t0[C][q] - t0[cur][q];
    }
    double tmp = sqrt(SQR(cA[0]) + SQR(cA[1]) + SQR(cA[2])) / sqrt(SQR(cC[0]) + SQR(cC[1]) + SQR(cC[2]));
    for (int q = 0; q < 3; q++) t0[A][q] = t0[cur][q] + tmp * cC[q];
    tmp = 1.0 / tmp;
    for (int q = 0; q < 3; q++) t0[C][q] = t0[cur][q] + tmp * cA[q];
    for (int j = 0; j < nb; j++) eobond[j] = enbond[j] = CH_BOND_E(j);
    for (int j = 0; j < nbe; j++) eobend[j] = enbend[j] = CH_BEND_E(j);
  }
  for (int jjj = 0; jjj < trialmove; jjj++) {
    int choise;
    if (nbe == 1)
      choise = cd(d) < 0.3 ? 1 : 2;
    else if (jjj < 20)
      choise = 3;
    else {
      double tmp = cd(d);
      choise = tmp < 0.2 ? 1 : (tmp < 0.6 ? 2 : 3);
    }
    for (int j = 0; j < nb; j++) enbond[j] = eobond[j];
    for (int j = 0; j < nbe; j++) enbend[j] = eobend[j];
    double newbond = 0.0, newbend = 0.0;
    if (choise == 1) {
      if (nb > 0) {
        int jj = (int)(cd(d) * (double)nb);
        if (jj >= nb) jj = nb - 1; /* u = 1 (closed interval) would index past the list */
        if (c->bond[st->bonds[jj]][2] != CH_FIXED_BOND) {
          int jo = st->branch[jj];
          double store[3] = {t0[jo][0], t0[jo][1], t0[jo][2]};
          double vec[3] = {t0[jo][0] - t0[cur][0], t0[jo][1] - t0[cur][1], t0[jo][2] - t0[cur][2]};
          double blo = sqrt(SQR(vec[0]) + SQR(vec[1]) + SQR(vec[2]));
          double bln = blo + (2.0 * cd(d) - 1.0) * c->max_bond[jo];
          for (int q = 0; q < 3; q++) t0[jo][q] = t0[cur][q] + (bln / blo) * vec[q];
          enbond[jj] = CH_BOND_E(jj);
          newbond += enbond[jj] - eobond[jj];
          if (cd(d) < SQR(bln / blo) * exp(-beta * (newbond + 0.0 + 0.0 + 0.0)))
            eobond[jj] = enbond[jj];
          else
            for (int q = 0; q < 3; q++) t0[jo][q] = store[q];
        }
      }
    } else if (choise == 2) {
      if (st->n_branch > 0) {
        int jj = (int)(cd(d) * (double)st->n_branch);
        if (jj >= st->n_branch) jj = st->n_branch - 1;
        int jo = st->branch[jj];
        double store[3] = {t0[jo][0], t0[jo][1], t0[jo][2]};
        double cord[1][3] = {{t0[jo][0] - t0[cur][0], t0[jo][1] - t0[cur][1], t0[jo][2] - t0[cur][2]}};
        double SinusOld = sin(ch_angle(t0[prev], t0[cur], t0[jo]));
        double dr[3] = {t0[cur][0] - t0[jo][0], t0[cur][1] - t0[jo][1], t0[cur][2] - t0[jo][2]};
        double perpen[3];
        ch_perpendicular(dr, pv, perpen);
        ch_rotate(perpen, cord, 1, (2.0 * cd(d) - 1.0) * c->max_bend[cur]);
        for (int q = 0; q < 3; q++) t0[jo][q] = t0[cur][q] + cord[0][q];
        double SinusNew = sin(ch_angle(t0[prev], t0[cur], t0[jo]));
        for (int j = 0; j < nbe; j++) {
          enbend[j] = CH_BEND_E(j);
          newbend += enbend[j] - eobend[j];
        }
        if (cd(d) < ((SinusNew / SinusOld) * exp(-beta * (newbend + 0.0 + 0.0 + 0.0)))) {
          for (int j = 0; j < nbe; j++) eobend[j] = enbend[j];
        } else
          for (int q = 0; q < 3; q++) t0[jo][q] = store[q];
      }
    } else {
      if (st->n_branch > 0) {
        int jj = (int)(cd(d) * (double)st->n_branch);
        if (jj >= st->n_branch) jj = st->n_branch - 1;
        int jo = st->branch[jj];
        double store[3] = {t0[jo][0], t0[jo][1], t0[jo][2]};
        double cord[1][3] = {{t0[jo][0] - t0[cur][0], t0[jo][1] - t0[cur][1], t0[jo][2] - t0[cur][2]}};
        ch_rotate(pv, cord, 1, (2.0 * cd(d) - 1.0) * c->max_cone[cur]);
        for (int q = 0; q < 3; q++) t0[jo][q] = t0[cur][q] + cord[0][q];
        for (int j = 0; j < nbe; j++) {
          enbend[j] = CH_BEND_E(j);
          newbend += enbend[j] - eobend[j];
        }
        if (cd(d) < exp(-beta * (newbend + 0.0 + 0.0 + 0.0))) {
          for (int j = 0; j < nbe; j++) eobend[j] = enbend[j];
        } else
          for (int q = 0; q < 3; q++) t0[jo][q] = store[q];
      }
    }
  }
#undef CH_BOND_E
#undef CH_BEND_E
  /* torsion biasing: rotations about prev_vec leave bonds and bends alone */
  double store[ORC_CHAIN_MAX][3];
  for (int jj = 0; jj < st->n_place; jj++)
    for (int q = 0; q < 3; q++) store[jj][q] = t0[st->place[jj]][q];
  static double bf[CH_MAX_TORSION_TRIALS], angle[CH_MAX_TORSION_TRIALS];
  static int ov[CH_MAX_TORSION_TRIALS];
  for (int iu = 0; iu < k; iu++) {
    cd_sub(d, ch_tag(step, 2 + iu));
    double cord[ORC_CHAIN_MAX][3];
    for (int kk = 0; kk < kt; kk++) {
      for (int jj = 0; jj < st->n_place; jj++) {
        int a = st->place[jj];
        for (int q = 0; q < 3; q++) {
          w->tpos[iu][a][q] = store[jj][q];
          cord[jj][q] = w->tpos[iu][a][q] - w->tpos[iu][cur][q];
        }
      }
      angle[kk] = (2.0 * cd(d) - 1.0) * M_PI;
      ch_rotate(pv, cord, st->n_place, angle[kk]);
      for (int jj = 0; jj < st->n_place; jj++)
        for (int q = 0; q < 3; q++) w->tpos[iu][st->place[jj]][q] = cord[jj][q] + w->tpos[iu][cur][q];
      double newtrs = ch_sum_torsions(c, w, st, iu);
      bf[kk] = -beta * newtrs;
      ov[kk] = 0;
    }
    double sum = 0.0;
    for (int i = 0; i < k; i++)
      if (!ov[i]) sum += exp(bf[i]);
    RT[iu] = sum / (double)k;
    int sel = st->n_tors > 0 ? ch_select(bf, ov, kt, cd(d)) : (int)(cd(d) * (double)kt);
    if (sel >= kt) sel = kt - 1;
    for (int jj = 0; jj < st->n_place; jj++) {
      int a = st->place[jj];
      for (int q = 0; q < 3; q++) {
        w->tpos[iu][a][q] = store[jj][q];
        cord[jj][q] = w->tpos[iu][a][q] - w->tpos[iu][cur][q];
      }
    }
    ch_rotate(pv, cord, st->n_place, angle[sel]);
    for (int jj = 0; jj < st->n_place; jj++)
      for (int q = 0; q < 3; q++) w->tpos[iu][st->place[jj]][q] = cord[jj][q] + w->tpos[iu][cur][q];
  }
}

static void chain_insertion(const orc_system *s, const orc_chain *c, chain_work *w, cdraw_t *d,
                            double rec[ORC_CREC_N], double *xyz) {
  const double crit = ENERGY_OVERLAP_CRITERIA, beta = s->beta;
  const int k_fb = c->k_fb, k = c->k_trial, n = c->n_atoms;
  for (int i = 0; i < ORC_CREC_N; i++) rec[i] = 0.0;
  /* HandleFirstBead (src/cbmc.c:623-795) */
  double fbpos[64][3], bf[64], ev[64], ei[64], eh[64];
  int ov[64];
  int OVERLAP = 1;
  const int type0 = c->type[c->start];
  for (int i = 0; i < k_fb; i++) {
    double sv[3];
    sv[2] = cd(d);
    sv[1] = cd(d);
    sv[0] = cd(d);
    fbpos[i][0] = s->box[AX] * sv[0] + s->box[BX] * sv[1] + s->box[CX] * sv[2];
    fbpos[i][1] = s->box[AY] * sv[0] + s->box[BY] * sv[1] + s->box[CY] * sv[2];
    fbpos[i][2] = s->box[AZ] * sv[0] + s->box[BZ] * sv[1] + s->box[CZ] * sv[2];
    double Uvdw = orc_vdw_energy_at(s, type0, fbpos[i], 1.0);
    if (Uvdw >= crit) {
      ov[i] = 1;
    } else {
      ov[i] = 0;
      OVERLAP = 0;
      ev[i] = Uvdw;
      bf[i] = -beta * (Uvdw + 0.0 + 0.0);
    }
  }
  if (OVERLAP) {
    rec[ORC_CREC_OVERLAP] = 1.0;
    rec[ORC_CREC_STAGE] = 0.0;
    return;
  }
  double W1 = 0.0;
  for (int i = 0; i < k_fb; i++)
    if (!ov[i]) W1 += exp(bf[i]);
  int sel = ch_select(bf, ov, k_fb, cd(d));
  W1 /= (double)k_fb;
  rec[ORC_CREC_W1] = W1;
  if (W1 < MINIMUM_ROSENBLUTH_FACTOR) {
    rec[ORC_CREC_OVERLAP] = 1.0;
    rec[ORC_CREC_STAGE] = 1.0;
    return;
  }
  double W = W1, UHost = ev[sel], UBond = 0.0, UBend = 0.0, UTors = 0.0, UIntra = 0.0;
  double pos[ORC_CHAIN_MAX][3];
  char placed[ORC_CHAIN_MAX] = {0};
  for (int q = 0; q < 3; q++) pos[c->start][q] = fbpos[sel][q];
  placed[c->start] = 1;
  /* Rosen (src/cbmc.c:2762-2849) */
  for (int si = 0; si < w->n_steps; si++) {
    const chain_step *st = &w->steps[si];
    const int step = si + 1;
    for (int iu = 0; iu < k; iu++)
      for (int a = 0; a < n; a++)
        if (placed[a])
          for (int q = 0; q < 3; q++) w->tpos[iu][a][q] = pos[a][q];
    double RT[64], UBondT[64], UBendT[64], UTorsT[64];
    for (int iu = 0; iu < k; iu++) {
      RT[iu] = 1.0;
      UBondT[iu] = UBendT[iu] = UTorsT[iu] = 0.0;
    }
    if (st->n_prev == 0) {
      /* GenerateTrialOrientationsSimpleSphere(FALSE), flexible group */
      cd_sub(d, ch_tag(step, 0));
      double bl[ORC_CHAIN_MAX];
      for (int j = 0; j < st->n_bonds; j++) bl[j] = ch_bond_length(d, beta, c->bond[st->bonds[j]][2], w->par.bond[st->bonds[j]]);
      const int a = st->branch[0];
      for (int iu = 0; iu < k; iu++) {
        cd_sub(d, ch_tag(step, 2 + iu));
        double vec[3];
        ch_sphere(d, vec);
        for (int q = 0; q < 3; q++) w->tpos[iu][a][q] = w->tpos[0][st->cur][q] + vec[q] * bl[0];
      }
      for (int iu = 0; iu < k; iu++) {
        int b = st->bonds[0];
        UBondT[iu] = ch_bond_u(c->bond[b][2], w->par.bond[b], w->tpos[iu][c->bond[b][0]], w->tpos[iu][c->bond[b][1]]);
      }
    } else {
      ch_mc_scheme(s, c, w, st, step, d, RT);
      for (int iu = 0; iu < k; iu++) {
        UBondT[iu] = ch_sum_bonds(c, w, st, iu);
        UBendT[iu] = ch_sum_bends(c, w, st, iu);
        UTorsT[iu] = ch_sum_torsions(c, w, st, iu);
      }
    }
    /* ComputeExternalEnergies (src/cbmc.c:1154-1478), LJ biasing */
    int any = 0;
    for (int iu = 0; iu < k; iu++) {
      int trial_overlap = 0;
      double ehost = 0.0;
      for (int j = 0; j < st->n_place; j++) {
        int a = st->place[j];
        double u = orc_vdw_energy_at(s, c->type[a], w->tpos[iu][a], 1.0);
        if (u >= crit) {
          trial_overlap = 1;
          break;
        }
        ehost += u;
      }
      if (trial_overlap) {
        bf[iu] = 0.0;
        ov[iu] = 1;
        continue;
      }
      double intra = 0.0;
      for (int i = 0; i < st->n_intra; i++) {
        int p = st->intra[i];
        int A = c->intra[p][0], B = c->intra[p][1];
        double rr = SQR(w->tpos[iu][A][0] - w->tpos[iu][B][0]) + SQR(w->tpos[iu][A][1] - w->tpos[iu][B][1]) +
                    SQR(w->tpos[iu][A][2] - w->tpos[iu][B][2]);
        if (rr < s->rc_vdw2) intra += w->par.scal[p] * lj_value(w->par.lj[p][0], w->par.lj[p][1], w->par.lj[p][2], rr);
      }
      eh[iu] = ehost;
      ei[iu] = intra;
      bf[iu] = -beta * (ehost + 0.0 + 0.0 + intra);
      ov[iu] = 0;
      any = 1;
    }
    if (!any) {
      rec[ORC_CREC_OVERLAP] = 1.0;
      rec[ORC_CREC_STAGE] = 2.0;
      rec[ORC_CREC_STEP] = step;
      return;
    }
    double weight = 0.0;
    for (int iu = 0; iu < k; iu++)
      if (!ov[iu]) weight += exp(bf[iu]);
    weight /= (double)k;
    cd_sub(d, ch_tag(step, 1));
    int iwalk = ch_select(bf, ov, k, cd(d));
    W *= weight * RT[iwalk];
    if (W < MINIMUM_ROSENBLUTH_FACTOR) {
      rec[ORC_CREC_OVERLAP] = 1.0;
      rec[ORC_CREC_STAGE] = 3.0;
      rec[ORC_CREC_STEP] = step;
      return;
    }
    UBond += UBondT[iwalk];
    UBend += UBendT[iwalk];
    UTors += UTorsT[iwalk];
    UIntra += ei[iwalk];
    UHost += eh[iwalk];
    for (int j = 0; j < st->n_place; j++) {
      int a = st->place[j];
      for (int q = 0; q < 3; q++) pos[a][q] = w->tpos[iwalk][a][q];
      placed[a] = 1;
    }
  }
  if (xyz)
    for (int a = 0; a < n; a++)
      for (int q = 0; q < 3; q++) xyz[3 * a + q] = pos[a][q];
  /* the GrowMolecule corrections (src/cbmc.c:3140-3177) and the tail term
   * (src/mc_moves.c:7888-7891) are exp(-beta * 0) = 1 here */
  for (int a = 0; a < n; a++)
    if (orc_blocked_pocket(s, pos[a])) {
      rec[ORC_CREC_STAGE] = 5.0;
      rec[ORC_CREC_UHOSTVDW] = UHost;
      return;
    }
  rec[ORC_CREC_W] = W;
  rec[ORC_CREC_DU] = UBond + UBend + UTors + UIntra + UHost; /* src/mc_moves.c:7912-7925 */
  rec[ORC_CREC_UHOSTVDW] = UHost;
  rec[ORC_CREC_UINTRA] = UIntra;
  rec[ORC_CREC_UBOND] = UBond;
  rec[ORC_CREC_UBEND] = UBend;
  rec[ORC_CREC_UTORSION] = UTors;
  rec[ORC_CREC_STAGE] = 4.0;
}

static chain_work *chain_prepare(const orc_system *s, const orc_chain *c) {
  chain_work *w = (chain_work *)malloc(sizeof(chain_work));
  w->n_steps = chain_steps(c, w->steps);
  chain_params(s, c, &w->par);
  return w;
}

void orc_chain_insertion_stream(const orc_system *s, const orc_chain *c, uint64_t seed, int64_t insertion,
                                double rec[ORC_CREC_N], double *xyz) {
  chain_work *w = chain_prepare(s, c);
  cdraw_t d = {NULL, orc_stream_key(seed, (uint64_t)insertion), 0, 0};
  d.sub = d.key;
  chain_insertion(s, c, w, &d, rec, xyz);
  free(w);
}

void orc_chain_insertion_mt(const orc_system *s, const orc_chain *c, orc_mt64 *g, double rec[ORC_CREC_N],
                            double *xyz) {
  chain_work *w = chain_prepare(s, c);
  cdraw_t d = {g, 0, 0, 0};
  chain_insertion(s, c, w, &d, rec, xyz);
  free(w);
}

void orc_chain_run_stream(const orc_system *s, const orc_chain *c, uint64_t seed, int64_t first, int64_t n,
                          double *rec, double *xyz) {
  chain_work *w = chain_prepare(s, c);
  for (int64_t i = 0; i < n; i++) {
    cdraw_t d = {NULL, orc_stream_key(seed, (uint64_t)(first + i)), 0, 0};
    d.sub = d.key;
    chain_insertion(s, c, w, &d, rec + ORC_CREC_N * i, xyz ? xyz + (size_t)3 * c->n_atoms * i : NULL);
  }
  free(w);
}

void orc_chain_run_mt(const orc_system *s, const orc_chain *c, uint64_t seed, int64_t n, double *rec,
                      double *xyz) {
  chain_work *w = chain_prepare(s, c);
  orc_mt64 g;
  orc_mt64_seed(&g, seed);
  for (int64_t i = 0; i < n; i++) {
    cdraw_t d = {&g, 0, 0, 0};
    chain_insertion(s, c, w, &d, rec + ORC_CREC_N * i, xyz ? xyz + (size_t)3 * c->n_atoms * i : NULL);
  }
  free(w);
}

/* ===================================================================== */
/* Pore-size distribution (SamplePoreSizeDistribution, src/sample.c:3444-  */
/* 3560; CheckSphereOverlap :3376-3440; ComputePoreSizeDistribution        */
/* src/monte_carlo.c:1515-1540)                                            */
/* ===================================================================== */

/* CheckSphereOverlap: 1 when pos is blocked or closer than
 * 0.5 factor sigma_jj to a framework atom j (the first such atom ends the
 * search); otherwise *radius = min_j |pos - r_j| - 0.5 factor sigma_jj */
static int psd_overlap(const orc_system *s, double factor, const double pos[3], double *radius) {
  *radius = DBL_MAX;
  if (orc_blocked_pocket(s, pos)) return 1;
  for (int j = 0; j < s->n; j++) {
    double dr[3] = {pos[0] - s->x[j], pos[1] - s->y[j], pos[2] - s->z[j]};
    orc_min_image(s->box, s->inv, s->bc, dr);
    double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
    if (rr < SQR(0.5 * factor * s->sig[j])) return 1;
    double r = sqrt(rr) - 0.5 * factor * s->sig[j];
    if (r < *radius) *radius = r;
  }
  return 0;
}

static void psd_point(const orc_system *s, double u0, double u1, double u2, double p[3]) {
  /* ConvertFromABCtoXYZ (src/potentials.c:5548-5556) */
  p[0] = s->box[AX] * u0 + s->box[BX] * u1 + s->box[CX] * u2;
  p[1] = s->box[AY] * u0 + s->box[BY] * u1 + s->box[CY] * u2;
  p[2] = s->box[AZ] * u0 + s->box[BZ] * u1 + s->box[CZ] * u2;
}

/* LargestRadius of cycle `i` (-10 when A overlaps or no sphere holds A):
 * point A from counters 0..2 of the cycle's stream, centre B_j from counters
 * 3+3j..5+3j, each in x, y, z order */
double orc_psd_largest(const orc_system *s, double factor, int64_t n_b, uint64_t seed, int64_t i) {
  uint64_t key = orc_stream_key(seed, (uint64_t)i);
  double a[3], radius;
  psd_point(s, orc_stream_uniform(key, 0), orc_stream_uniform(key, 1), orc_stream_uniform(key, 2), a);
  if (psd_overlap(s, factor, a, &radius)) return -10.0;
  double largest = -10.0;
  for (int64_t j = 0; j < n_b; j++) {
    double b[3];
    psd_point(s, orc_stream_uniform(key, (uint64_t)(3 + 3 * j)), orc_stream_uniform(key, (uint64_t)(4 + 3 * j)),
              orc_stream_uniform(key, (uint64_t)(5 + 3 * j)), b);
    if (psd_overlap(s, factor, b, &radius)) continue;
    double dr[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    orc_min_image(s->box, s->inv, s->bc, dr);
    double rr = SQR(dr[0]) + SQR(dr[1]) + SQR(dr[2]);
    double r = sqrt(rr);
    if (r > radius) continue;
    if (radius > largest) largest = radius;
  }
  return largest;
}

/* PoreSizeDistributionHistogram after cycles [first, first+n): hist[k] (k <
 * bins) counts the cycles whose largest radius falls in bin k or above, as
 * the reference adds 1 to every bin up to the cycle's (a radius beyond the
 * range adds nothing) */
void orc_psd_histogram(const orc_system *s, double factor, int64_t n_b, uint64_t seed, int64_t first,
                       int64_t n, int bins, double range, double *hist) {
  double deltaR = range / bins;
  for (int64_t i = first; i < first + n; i++) {
    double largest = orc_psd_largest(s, factor, n_b, seed, i);
    int index = (int)(largest / deltaR);
    if ((index >= 0) && (index < bins))
      for (int k = 0; k <= index; k++) hist[k] += 1.0;
  }
}
