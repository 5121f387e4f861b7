// rhg_common.h -- types and inline math shared by the host runtime and the
// gfx950 kernels of librhg.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef RHG_HD
#define RHG_HD __host__ __device__ __forceinline__
#endif

namespace rhg {

constexpr double kOverlap = 1.0e7;          // EnergyOverlapCriteria (src/input.c:367)
constexpr double kMinRosenbluth = 1.0e-150; // MinimumRosenbluthFactor (src/input.c:368)
constexpr double kInvMax53 = 1.0 / 9007199254740991.0;
constexpr int kMaxSites = 8;
constexpr int kMaxTrials = 64;

enum { AX = 0, AY, AZ, BX, BY, BZ, CX, CY, CZ };

// Box and InverseBox in RASPA2's field order; bc 1 rectangular, 2 triclinic.
struct Cell {
  double box[9];
  double inv[9];
  int bc;
  int upper; // box and inverse have exact zeros at ay, az, bz
};

// Framework atom as seen by one probe pseudo-atom, FP64 arithmetic.
// c4 = 4*eps_ij (exact scaling of PotentialParms[0]), s2 = sigma_ij^2,
// shift = PotentialParms[2], q = framework charge.
struct __attribute__((aligned(16))) AtomRec64 {
  double x, y, z, c4, s2, shift, q, pad;
};

// Same atom for the mixed kernels: FP64 position, FP32 pair parameters.
struct __attribute__((aligned(8))) AtomRecMx {
  double x, y, z;
  float c4, s2, shift, q;
};

// ------------------------------------------------------------- rounding
// (double)NINT(x) of src/utils.h:60: round half away from zero, via the
// same two steps (add/subtract 0.5, then truncate).
RHG_HD double nint_d(double x) {
  double t = (x >= 0.0) ? (x + 0.5) : (x - 0.5);
  return trunc(t);
}

// ApplyBoundaryCondition (src/potentials.h:232-271) in the reference's
// operation order, no contraction.  Terms whose matrix entry is an exact
// zero are skipped when cell.upper is set; that changes at most the sign of
// a zero, never r^2.
RHG_HD void min_image_exact(const Cell &c, double &dx, double &dy, double &dz) {
#pragma clang fp contract(off)
  if (c.bc == 1) {
    dx -= c.box[AX] * nint_d(dx * c.inv[AX]);
    dy -= c.box[BY] * nint_d(dy * c.inv[BY]);
    dz -= c.box[CZ] * nint_d(dz * c.inv[CZ]);
  } else if (c.upper) {
    double sx = c.inv[AX] * dx + c.inv[BX] * dy + c.inv[CX] * dz;
    double sy = c.inv[BY] * dy + c.inv[CY] * dz;
    double sz = c.inv[CZ] * dz;
    double tx = sx - nint_d(sx);
    double ty = sy - nint_d(sy);
    double tz = sz - nint_d(sz);
    dx = c.box[AX] * tx + c.box[BX] * ty + c.box[CX] * tz;
    dy = c.box[BY] * ty + c.box[CY] * tz;
    dz = c.box[CZ] * tz;
  } else {
    double sx = c.inv[AX] * dx + c.inv[BX] * dy + c.inv[CX] * dz;
    double sy = c.inv[AY] * dx + c.inv[BY] * dy + c.inv[CY] * dz;
    double sz = c.inv[AZ] * dx + c.inv[BZ] * dy + c.inv[CZ] * dz;
    double tx = sx - nint_d(sx);
    double ty = sy - nint_d(sy);
    double tz = sz - nint_d(sz);
    dx = c.box[AX] * tx + c.box[BX] * ty + c.box[CX] * tz;
    dy = c.box[AY] * tx + c.box[BY] * ty + c.box[CY] * tz;
    dz = c.box[AZ] * tx + c.box[BZ] * ty + c.box[CZ] * tz;
  }
}

// LENNARD_JONES branch of PotentialValue (src/potentials.c:5867-5877)
// with c4 = 4*arg1, s2 = arg2 = sigma^2, in the reference's order.
RHG_HD double lj_exact(double c4, double s2, double shift, double rr) {
#pragma clang fp contract(off)
  double t = s2 / rr;
  double rri3 = (t * t) * t;
  return c4 * (rri3 * (rri3 - 1.0)) - shift;
}

// ------------------------------------------------------ counter stream
// SplitMix64 output function (Steele, Lea & Flood 2014) over key +
// (k+1)*golden gamma; uniform = RandomNumber()'s closed-interval mapping
// (src/utils.c:378-381).
RHG_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
RHG_HD uint64_t stream_key(uint64_t seed, uint64_t insertion) {
  return mix64(seed ^ mix64(insertion + 0x632BE59BD9B4E019ULL));
}
RHG_HD double stream_uniform(uint64_t key, uint64_t counter) {
  uint64_t x = mix64(key + (counter + 1ULL) * 0x9E3779B97F4A7C15ULL);
  return (double)(x >> 11) * kInvMax53;
}
// Sub-stream `tag` of an insertion's key (flexible-chain growth draws one
// sub-stream per purpose so that parallel kernels know where their draws
// start): tag = (step << 32) | j.
RHG_HD uint64_t chain_substream(uint64_t key, uint64_t tag) {
  return mix64(key ^ mix64(tag + 0x2545F4914F6CDD1DULL));
}
RHG_HD uint64_t chain_tag(int step, int j) { return ((uint64_t)step << 32) | (uint64_t)j; }

// ------------------------------------------------------ portable sincos
// sin and cos of x in [0, 2*pi] (the Shoemake angles 2*pi*u), written so
// that host and device produce identical bits: Cody-Waite reduction by
// pi/2 with a two-part constant, then Taylor series of sin and cos on
// |r| <= pi/4 through r^21 / r^20 (truncation error < 2e-24), Horner form,
// no contraction.  Accuracy is within a few ulp of the correctly rounded
// value.
RHG_HD void sincos_portable(double x, double *s, double *c) {
#pragma clang fp contract(off)
  const double two_over_pi = 0.63661977236758138243;
  const double pio2_hi = 1.57079632673412561417e+00; // first 33 bits of pi/2
  const double pio2_lo = 6.07710050650619224932e-11; // pi/2 - pio2_hi
  double kd = trunc(x * two_over_pi + 0.5);
  int k = (int)kd;
  double r = (x - kd * pio2_hi) - kd * pio2_lo;
  double r2 = r * r;
  // 1/(2n+1)! and 1/(2n)!
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

// RandomArrayRotationMatrix (src/utils.c:835-877) applied to one vector,
// given the three Shoemake uniforms; trig from sincos_portable.
struct Rot {
  double R[3][3];
};
RHG_HD Rot shoemake(double X0, double u1, double u2) {
#pragma clang fp contract(off)
  Rot o;
  double Y1 = 2.0 * M_PI * u1;
  double Y2 = 2.0 * M_PI * u2;
  double R1 = sqrt(1.0 - X0);
  double R2 = sqrt(X0);
  double sY1, cY1, sY2, cY2;
  sincos_portable(Y1, &sY1, &cY1);
  sincos_portable(Y2, &sY2, &cY2);
  double U0 = cY2 * R2;
  double U1 = sY1 * R1;
  double U2 = cY1 * R1;
  double U3 = sY2 * R2;
  double COEFI = 2.0 * U0 * U0 - 1.0;
  double COEFUU = 2.0;
  double COEFE = 2.0 * U0;
  o.R[0][0] = COEFI + COEFUU * U1 * U1;
  o.R[1][1] = COEFI + COEFUU * U2 * U2;
  o.R[2][2] = COEFI + COEFUU * U3 * U3;
  o.R[1][2] = COEFUU * U2 * U3 - COEFE * U1;
  o.R[2][0] = COEFUU * U3 * U1 - COEFE * U2;
  o.R[0][1] = COEFUU * U1 * U2 - COEFE * U3;
  o.R[2][1] = COEFUU * U3 * U2 + COEFE * U1;
  o.R[0][2] = COEFUU * U1 * U3 + COEFE * U2;
  o.R[1][0] = COEFUU * U2 * U1 + COEFE * U3;
  return o;
}
RHG_HD void rotate(const Rot &o, const double v[3], double out[3]) {
#pragma clang fp contract(off)
  out[0] = v[0] * o.R[0][0] + v[1] * o.R[0][1] + v[2] * o.R[0][2];
  out[1] = v[0] * o.R[1][0] + v[1] * o.R[1][1] + v[2] * o.R[1][2];
  out[2] = v[0] * o.R[2][0] + v[1] * o.R[2][1] + v[2] * o.R[2][2];
}

// Molecule description on the device.
struct MolDev {
  int n_sites, start, k_fb, k_or, n_place;
  int place[kMaxSites];       // BeadsToBePlaced in group order
  double body[kMaxSites][3];  // Components[c].Positions
  double q[kMaxSites];        // site charges (0 for uncharged)
  int charged[kMaxSites];     // HasCharges
  // storage of the placed beads, grouped by pseudo-atom type so that each
  // type is one site list: for placed bead p, the first storage slot of its
  // group, the group size and p's index within the group
  int g_start[kMaxSites], g_size[kMaxSites], g_b[kMaxSites];
  // the same over the charged placed beads (h_start < 0: p is uncharged)
  int h_start[kMaxSites], h_size[kMaxSites], h_b[kMaxSites];
};

} // namespace rhg
