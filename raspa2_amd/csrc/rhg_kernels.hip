// rhg_kernels.hip -- gfx950 kernels of the host-guest energy engine.
//
// Site-energy kernels give every lane one trial site and stream the
// framework through the scalar unit: the atom index of the inner loop is
// wave-uniform, so each atom record is fetched once per wave with scalar
// loads into SGPRs and broadcast to all 64 lanes as an instruction operand
// (no LDS traffic, no bank conflicts, no VGPRs for the framework).
#include <float.h>
#include <cstring>
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "rhg_kernels.h"

namespace rhg {

// erfc(x) in FP32 for the mixed real-space Ewald terms:
// exp(-x^2) P(t), t = 1 / (1 + x/2), with P of degree 8 fitted to
// erfcx(x) = exp(x^2) erfc(x) on [0, 5] (raspa2_amd/csrc/fit_erfc.py; relative
// error 3.5e-8 in exact arithmetic).  exp(-x^2) is one exp2 of the rounded
// -x^2 log2(e): the rounding of that exponent grows with x^2, so the relative
// error of the result reaches 1.1e-6 at x = 3.2 (2.6e-6 at 5), while the
// absolute error stays below 4e-7 over [0, 5]; a pair term enters a site's
// sum with its absolute error, which is what the Coulomb tolerances bound.
// Larger arguments (only for Ewald precisions beyond 1e-10) use the library
// erfcf.
__device__ __forceinline__ float erfcx_poly(float t) {
  float p = -0.040972326f;
  p = fmaf(p, t, 0.292736441f);
  p = fmaf(p, t, -0.786886334f);
  p = fmaf(p, t, 0.885848165f);
  p = fmaf(p, t, -0.305710703f);
  p = fmaf(p, t, 0.427213341f);
  p = fmaf(p, t, 0.240593985f);
  p = fmaf(p, t, 0.287479192f);
  p = fmaf(p, t, -0.000301792286f);
  return p;
}
__device__ __forceinline__ float erfc_fit(float x) {
  const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, x, 1.0f));
  const float p = erfcx_poly(t);
  // exp(-x^2) = 2^(-x^2 log2 e)
  return p * __builtin_amdgcn_exp2f((x * x) * -1.44269502f);
}

// The same fit as erfc(alpha r)/r from r^2 and 1/r (the pair loops have
// both): t = 1 / (1 + (alpha/2 r^2) (1/r)) and exp(-alpha^2 r^2) as one exp2
// of r^2 ke, with the constants of coulomb_consts (two instructions fewer
// than forming r and alpha r first).
struct CoulConsts {
  float bh; // alpha / 2
  float ke; // -alpha^2 log2(e)
};
__device__ __forceinline__ CoulConsts coulomb_consts(double alpha) {
  const float a = (float)alpha;
  return CoulConsts{0.5f * a, -(a * a) * 1.44269502f};
}
__device__ __forceinline__ float erfc_over_r(float r2, float irs, CoulConsts k) {
  const float t = __builtin_amdgcn_rcpf(fmaf(k.bh * r2, irs, 1.0f));
  return (erfcx_poly(t) * __builtin_amdgcn_exp2f(r2 * k.ke)) * irs;
}
__device__ __forceinline__ float erfc_mixed(float x) {
  if (x > 5.0f) return erfcf(x);
  return erfc_fit(x);
}

// ====================================================================
// minimum image variants
// ====================================================================
// BCK 1: rectangular, 2: upper-triangular triclinic, 3: general triclinic
template <int BCK>
__device__ __forceinline__ void mi_exact(const Cell &c, double &dx, double &dy, double &dz) {
#pragma clang fp contract(off)
  if (BCK == 1) {
    dx -= c.box[AX] * nint_d(dx * c.inv[AX]);
    dy -= c.box[BY] * nint_d(dy * c.inv[BY]);
    dz -= c.box[CZ] * nint_d(dz * c.inv[CZ]);
  } else if (BCK == 2) {
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

// Fast minimum image for the mixed kernels: round-to-nearest-even and FMA.
// Differs from NINT only when a coordinate sits exactly half a box away,
// where both images are equidistant.
template <int BCK>
__device__ __forceinline__ void mi_fast(const Cell &c, double &dx, double &dy, double &dz) {
  if (BCK == 1) {
    dx = fma(-c.box[AX], rint(dx * c.inv[AX]), dx);
    dy = fma(-c.box[BY], rint(dy * c.inv[BY]), dy);
    dz = fma(-c.box[CZ], rint(dz * c.inv[CZ]), dz);
  } else if (BCK == 2) {
    double sx = fma(c.inv[AX], dx, fma(c.inv[BX], dy, c.inv[CX] * dz));
    double sy = fma(c.inv[BY], dy, c.inv[CY] * dz);
    double sz = c.inv[CZ] * dz;
    double tx = sx - rint(sx), ty = sy - rint(sy), tz = sz - rint(sz);
    dx = fma(c.box[AX], tx, fma(c.box[BX], ty, c.box[CX] * tz));
    dy = fma(c.box[BY], ty, c.box[CY] * tz);
    dz = c.box[CZ] * tz;
  } else {
    double sx = fma(c.inv[AX], dx, fma(c.inv[BX], dy, c.inv[CX] * dz));
    double sy = fma(c.inv[AY], dx, fma(c.inv[BY], dy, c.inv[CY] * dz));
    double sz = fma(c.inv[AZ], dx, fma(c.inv[BZ], dy, c.inv[CZ] * dz));
    double tx = sx - rint(sx), ty = sy - rint(sy), tz = sz - rint(sz);
    dx = fma(c.box[AX], tx, fma(c.box[BX], ty, c.box[CX] * tz));
    dy = fma(c.box[AY], tx, fma(c.box[BY], ty, c.box[CY] * tz));
    dz = fma(c.box[AZ], tx, fma(c.box[BZ], ty, c.box[CZ] * tz));
  }
}

// Site i of a row-layout list: the rows of one insertion are visited
// together (i = insertion * rows + row), so neighbouring lanes hold nearby
// positions, which binning aggregates.
__device__ __forceinline__ int64_t site_phys(const SiteList &s, int64_t i) {
  if (s.in_stride <= 0) return i;
  // list indices stay below 2^31: 32-bit division
  const uint32_t rows = (uint32_t)s.count_mult, u = (uint32_t)i;
  return (int64_t)(u % rows) * s.in_stride + (int64_t)(u / rows);
}

__device__ __forceinline__ int64_t site_count(const SiteList &s) {
  return s.n_dev ? (int64_t)(*s.n_dev) * s.count_mult : s.n;
}

// Exact FP64 energies of one site in reference order (all atoms, index order).
template <int BCK>
__device__ __forceinline__ void exact_site(const Cell &cell, const AtomRec64 *__restrict__ a,
                                           int n_atoms, double rc2v, double rc2c, double alpha,
                                           double cqa, int do_vdw, int do_cc, double px, double py,
                                           double pz, double &uv, double &uc) {
#pragma clang fp contract(off)
  uv = 0.0;
  uc = 0.0;
  for (int j = 0; j < n_atoms; ++j) {
    const AtomRec64 r = a[j];
    double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    mi_exact<BCK>(cell, dx, dy, dz);
    double rr = dx * dx + dy * dy + dz * dz;
    if (do_vdw && rr < rc2v) uv += lj_exact(r.c4, r.s2, r.shift, rr);
    if (do_cc && rr < rc2c) {
      double rd = sqrt(rr);
      uc += cqa * r.q * (erfc(alpha * rd) / rd);
    }
  }
}

// ====================================================================
// FP64 site energies: bit-identical to the sequential reference loop
// (src/framework_energy.c:151-170 and 258-275) for the LJ term.
// ====================================================================
template <int BCK, int VDW, int CC>
__global__ __launch_bounds__(256) void k_site_energy64(
    const Cell cell, const AtomRec64 *__restrict__ atoms, const int n_atoms, const double rc2v,
    const double rc2c, const double alpha, const double cqa, const SiteList s,
    double *__restrict__ out_v, double *__restrict__ out_c, uint8_t *__restrict__ out_o) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = site_count(s);
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;
  const bool live = i < n;
  const int64_t k = site_phys(s, live ? i : 0);
  const double px = s.x[k], py = s.y[k], pz = s.z[k];
  double uv = 0.0, uc = 0.0;
#pragma unroll 2
  for (int j = 0; j < n_atoms; ++j) {
    const AtomRec64 r = atoms[j];
    double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    mi_exact<BCK>(cell, dx, dy, dz);
    double rr = dx * dx + dy * dy + dz * dz;
    if (VDW) {
      if (rr < rc2v) uv += lj_exact(r.c4, r.s2, r.shift, rr);
    }
    if (CC) {
      if (rr < rc2c) {
        double rd = sqrt(rr);
        uc += cqa * r.q * (erfc(alpha * rd) / rd);
      }
    }
  }
  if (live) {
    const int64_t oi = k;
    if (out_v) out_v[oi] = uv;
    if (out_c) out_c[oi] = uc;
    if (out_o) out_o[oi] = (uint8_t)((VDW && uv >= kOverlap) || (CC && uc >= kOverlap));
  }
}

// ====================================================================
// Mixed site energies: FP64 differences and minimum image, FP32 pair
// terms, FP64 accumulation.  Sites whose energy lands within 1e-4 of the
// overlap threshold are recomputed with the exact FP64 loop so overlap
// flags equal the FP64 reference bit for bit.
// ====================================================================
template <int BCK, int VDW, int CC>
__global__ __launch_bounds__(256) void k_site_energy_mx(
    const Cell cell, const AtomRecMx *__restrict__ atoms, const AtomRec64 *__restrict__ atoms64,
    const int n_atoms, const double rc2v, const double rc2c, const double alpha, const double cqa,
    const SiteList s, double *__restrict__ out_v, double *__restrict__ out_c,
    uint8_t *__restrict__ out_o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = site_count(s);
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;
  const bool live = i < n;
  const int64_t k = site_phys(s, live ? i : 0);
  const double px = s.x[k], py = s.y[k], pz = s.z[k];
  const float alphaf = (float)alpha;
  const float cqaf = (float)cqa;
  double uv = 0.0, uc = 0.0;
#pragma unroll 4
  for (int j = 0; j < n_atoms; ++j) {
    const AtomRecMx r = atoms[j];
    double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    mi_fast<BCK>(cell, dx, dy, dz);
    const double rr = fma(dx, dx, fma(dy, dy, dz * dz));
    const float rrf = (float)rr;
    if (VDW) {
      const float ir = __builtin_amdgcn_rcpf(rrf);
      const float t = r.s2 * ir;
      const float r3 = t * t * t;
      const float u = fmaf(r.c4, fmaf(r3, r3, -r3), -r.shift);
      uv += (rr < rc2v) ? (double)u : 0.0;
    }
    if (CC) {
      if (rr < rc2c) {
        const float irs = __builtin_amdgcn_rsqf(rrf);
        const float rd = rrf * irs;
        const float e = erfc_mixed(alphaf * rd);
        uc += (double)((cqaf * r.q) * (e * irs));
      }
    }
  }
  const double band = 1.0e-4 * kOverlap;
  const bool near = live && ((VDW && fabs(uv - kOverlap) <= band) || (CC && fabs(uc - kOverlap) <= band));
  if (__any(near)) {
    if (near) {
      double ev, ec;
      exact_site<BCK>(cell, atoms64, n_atoms, rc2v, rc2c, alpha, cqa, VDW, CC, px, py, pz, ev, ec);
      if (VDW) uv = ev;
      if (CC) uc = ec;
    }
  }
  if (live) {
    const int64_t oi = k;
    if (out_v) out_v[oi] = uv;
    if (out_c) out_c[oi] = uc;
    if (out_o) out_o[oi] = (uint8_t)((VDW && uv >= kOverlap) || (CC && uc >= kOverlap));
  }
}

template <int BCK, int VDW, int CC>
static void site_energy_bck(const EnergyArgs &a, const SiteList &s, const EnergyOut &o,
                            hipStream_t st) {
  const int bs = 256;
  int64_t nb = (s.n + bs - 1) / bs;
  if (nb == 0) return;
  if (a.precision == 0)
    hipLaunchKernelGGL((k_site_energy64<BCK, VDW, CC>), dim3((unsigned)nb), dim3(bs), 0, st, a.cell,
                       a.a64, a.n_atoms, a.rc2_vdw, a.rc2_cc, a.alpha, a.cqa, s, o.u_vdw, o.u_cc,
                       o.overlap);
  else
    hipLaunchKernelGGL((k_site_energy_mx<BCK, VDW, CC>), dim3((unsigned)nb), dim3(bs), 0, st,
                       a.cell, a.amx, a.a64, a.n_atoms, a.rc2_vdw, a.rc2_cc, a.alpha, a.cqa, s,
                       o.u_vdw, o.u_cc, o.overlap);
}

template <int VDW, int CC>
static void site_energy_mode(const EnergyArgs &a, const SiteList &s, const EnergyOut &o,
                             hipStream_t st) {
  if (a.cell.bc == 1)
    site_energy_bck<1, VDW, CC>(a, s, o, st);
  else if (a.cell.upper)
    site_energy_bck<2, VDW, CC>(a, s, o, st);
  else
    site_energy_bck<3, VDW, CC>(a, s, o, st);
}

void launch_site_energy(const EnergyArgs &a, const SiteList &s, const EnergyOut &o,
                        hipStream_t st) {
  if (a.do_vdw && a.do_cc)
    site_energy_mode<1, 1>(a, s, o, st);
  else if (a.do_vdw)
    site_energy_mode<1, 0>(a, s, o, st);
  else if (a.do_cc)
    site_energy_mode<0, 1>(a, s, o, st);
}

// ====================================================================
// cell lists: neighbour lists of cells and tile-binned site energies
// ====================================================================
// periodic (torus) distance^2 between two points: NINT image in fractional
// coordinates, then the 27 neighbouring images (triclinic cells can put
// the nearest image one step away from the NINT one).
__device__ __forceinline__ double torus_dist2(const Cell &c, double dx, double dy, double dz) {
  double sx = c.inv[AX] * dx + c.inv[BX] * dy + c.inv[CX] * dz;
  double sy = c.inv[AY] * dx + c.inv[BY] * dy + c.inv[CY] * dz;
  double sz = c.inv[AZ] * dx + c.inv[BZ] * dy + c.inv[CZ] * dz;
  sx -= rint(sx);
  sy -= rint(sy);
  sz -= rint(sz);
  double best = DBL_MAX;
  const int r = (c.bc == 1) ? 0 : 1;
  for (int i = -r; i <= r; ++i)
    for (int j = -r; j <= r; ++j)
      for (int k = -r; k <= r; ++k) {
        double tx = sx + i, ty = sy + j, tz = sz + k;
        double x = c.box[AX] * tx + c.box[BX] * ty + c.box[CX] * tz;
        double y = c.box[AY] * tx + c.box[BY] * ty + c.box[CY] * tz;
        double z = c.box[AZ] * tx + c.box[BZ] * ty + c.box[CZ] * tz;
        best = fmin(best, x * x + y * y + z * z);
      }
  return best;
}

// one workgroup per cell; pass 0 counts, pass 1 writes in ascending order
// One wave per cell: the lanes test 64 atoms at a time against the cell
// centre and write the atoms within reach in ascending order (ballot
// compaction; no barriers).  pass 0 counts, pass 1 writes at nb_off[c].
__global__ __launch_bounds__(256) void k_cell_neighbours(const Cell cell, int nc0, int nc1, int nc2,
                                                         double reach2, const double *__restrict__ ax,
                                                         const double *__restrict__ ay,
                                                         const double *__restrict__ az, int n_atoms,
                                                         int32_t *__restrict__ counts,
                                                         const int32_t *__restrict__ nb_off,
                                                         int32_t *__restrict__ nb_idx, int pass) {
  const int c = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (c >= nc0 * nc1 * nc2) return;
  const int lane = threadIdx.x & 63;
  const int cz = c % nc2, cy = (c / nc2) % nc1, cx = c / (nc2 * nc1);
  const double fx = (cx + 0.5) / nc0, fy = (cy + 0.5) / nc1, fz = (cz + 0.5) / nc2;
  const double px = cell.box[AX] * fx + cell.box[BX] * fy + cell.box[CX] * fz;
  const double py = cell.box[AY] * fx + cell.box[BY] * fy + cell.box[CY] * fz;
  const double pz = cell.box[AZ] * fx + cell.box[BZ] * fy + cell.box[CZ] * fz;
  int base = pass ? nb_off[c] : 0;
  const unsigned long long lower = (1ULL << lane) - 1ULL;
  for (int j0 = 0; j0 < n_atoms; j0 += 64) {
    const int j = j0 + lane;
    const bool in = j < n_atoms && torus_dist2(cell, ax[j] - px, ay[j] - py, az[j] - pz) <= reach2;
    const unsigned long long m = __ballot(in);
    if (in && pass) nb_idx[base + __popcll(m & lower)] = j;
    base += __popcll(m);
  }
  if (lane == 0 && !pass) counts[c] = base;
}

void launch_cell_neighbours(const Cell &cell, int nc0, int nc1, int nc2, double reach,
                            const double *ax, const double *ay, const double *az, int n_atoms,
                            int32_t *counts, int32_t *nb_off, int32_t *nb_idx, int pass,
                            hipStream_t st) {
  const int ncell = nc0 * nc1 * nc2;
  hipLaunchKernelGGL(k_cell_neighbours, dim3((ncell + 3) / 4), dim3(256), 0, st, cell, nc0, nc1, nc2,
                     reach * reach, ax, ay, az, n_atoms, counts, nb_off, nb_idx, pass);
}

// single-workgroup exclusive scan of n ints into off[0..n]
__global__ __launch_bounds__(1024) void k_scan(const int32_t *__restrict__ in, int n,
                                               int32_t *__restrict__ off, int tiles) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int b = t * per, e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; ++i) s += tiles ? (in[i] + 63) / 64 : in[i];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < (int)blockDim.x; d <<= 1) {
    int v = (t >= d) ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t == 0) ? 0 : part[t - 1];
  for (int i = b; i < e; ++i) {
    off[i] = run;
    run += tiles ? (in[i] + 63) / 64 : in[i];
  }
  if (t == (int)blockDim.x - 1) off[n] = part[t];
}

void launch_scan_counts(const int32_t *counts, int n, int32_t *off, hipStream_t st) {
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, counts, n, off, 0);
}

__device__ __forceinline__ void cell_centre(const Cell &cell, const CellGrid &g, int c, double &px,
                                            double &py, double &pz) {
  const int cz = c % g.nc[2], cy = (c / g.nc[2]) % g.nc[1], cx = c / (g.nc[2] * g.nc[1]);
  const double fx = (cx + 0.5) / g.nc[0], fy = (cy + 0.5) / g.nc[1], fz = (cz + 0.5) / g.nc[2];
  px = cell.box[AX] * fx + cell.box[BX] * fy + cell.box[CX] * fz;
  py = cell.box[AY] * fx + cell.box[BY] * fy + cell.box[CY] * fz;
  pz = cell.box[AZ] * fx + cell.box[BZ] * fy + cell.box[CZ] * fz;
}

// One workgroup per cell over the cell's neighbour atoms (every atom within
// reach in torus distance).  The offset from the centre is wrapped to
// fractional components s in [-1/2, 1/2]; along a cell vector the shifted
// image s - sign(s) is at least (1 - |s|) W away (W: the perpendicular width)
// and s + sign(s) at least W, so each atom has at most 2 candidates per
// direction and 8 in all.  Candidates within reach are entries, atom-major in
// neighbour-list order, shifts in a fixed order, so the list order is fixed.
// One wave per cell: each lane takes one atom of the cell's neighbour list,
// finds its images within reach (at most two candidate shifts per axis) and
// the wave writes them in list order (a prefix sum of the per-lane counts,
// at most 8, from four ballots; no barriers).  pass 0 counts, pass 1 writes
// at off[c].  With box2 > 0 (rectangular cells) an image is kept only when
// its distance to the cell's box (half edges hx, hy, hz about the centre) is
// at most sqrt(box2): the images within the cut-off of some point of the
// cell, instead of all within the cut-off plus the half diagonal of its
// centre (the list holds about 7% fewer entries for a cube).
__global__ __launch_bounds__(256) void k_image_list(const Cell cell, const CellGrid g, double reach2,
                                                    double w0, double w1, double w2, double box2,
                                                    double hx, double hy, double hz,
                                                    const double *__restrict__ ax,
                                                    const double *__restrict__ ay,
                                                    const double *__restrict__ az,
                                                    int32_t *__restrict__ counts,
                                                    const int32_t *__restrict__ off,
                                                    float4 *__restrict__ img, int pass) {
  const int c = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (c >= g.ncell) return;
  const int lane = threadIdx.x & 63;
  const unsigned long long lower = (1ULL << lane) - 1ULL;
  double cx, cy, cz;
  cell_centre(cell, g, c, cx, cy, cz);
  int base = pass ? off[c] : 0;
  const double reach = sqrt(reach2);
  const int q0 = g.nb_off[c], cnt = g.nb_off[c + 1] - q0;
  for (int i0 = 0; i0 < cnt; i0 += 64) {
    const int i = i0 + lane;
    int j = 0, m = 0;
    double s[3] = {0.0, 0.0, 0.0};
    int cand[3][2] = {{0, 0}, {0, 0}, {0, 0}}, nc[3] = {1, 1, 1};
    if (i < cnt) {
      j = g.nb_idx[q0 + i];
      const double dx = ax[j] - cx, dy = ay[j] - cy, dz = az[j] - cz;
      s[0] = cell.inv[AX] * dx + cell.inv[BX] * dy + cell.inv[CX] * dz;
      s[1] = cell.inv[AY] * dx + cell.inv[BY] * dy + cell.inv[CY] * dz;
      s[2] = cell.inv[AZ] * dx + cell.inv[BZ] * dy + cell.inv[CZ] * dz;
      const double w[3] = {w0, w1, w2};
      for (int d = 0; d < 3; ++d) {
        s[d] -= rint(s[d]);
        if ((1.0 - fabs(s[d])) * w[d] <= reach) cand[d][nc[d]++] = s[d] > 0.0 ? -1 : 1;
      }
    }
    auto image = [&](int a, int b, int e, double &vx, double &vy, double &vz) {
      const double tx = s[0] + cand[0][a], ty = s[1] + cand[1][b], tz = s[2] + cand[2][e];
      vx = cell.box[AX] * tx + cell.box[BX] * ty + cell.box[CX] * tz;
      vy = cell.box[AY] * tx + cell.box[BY] * ty + cell.box[CY] * tz;
      vz = cell.box[AZ] * tx + cell.box[BZ] * ty + cell.box[CZ] * tz;
      if (box2 > 0.0) {
        const double ex = fmax(fabs(vx) - hx, 0.0), ey = fmax(fabs(vy) - hy, 0.0),
                     ez = fmax(fabs(vz) - hz, 0.0);
        return ex * ex + ey * ey + ez * ez <= box2;
      }
      return vx * vx + vy * vy + vz * vz <= reach2;
    };
    if (i < cnt)
      for (int a = 0; a < nc[0]; ++a)
        for (int b = 0; b < nc[1]; ++b)
          for (int e = 0; e < nc[2]; ++e) {
            double vx, vy, vz;
            m += image(a, b, e, vx, vy, vz) ? 1 : 0;
          }
    // exclusive prefix of m over the lanes and the wave's total (m <= 8)
    int pre = 0, tot = 0;
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
      const unsigned long long v = __ballot((m >> bit) & 1);
      pre += __popcll(v & lower) << bit;
      tot += __popcll(v) << bit;
    }
    int o = base + pre;
    if (pass && i < cnt)
      for (int a = 0; a < nc[0]; ++a)
        for (int b = 0; b < nc[1]; ++b)
          for (int e = 0; e < nc[2]; ++e) {
            double vx, vy, vz;
            if (image(a, b, e, vx, vy, vz))
              img[o++] = make_float4((float)vx, (float)vy, (float)vz, __int_as_float(j));
          }
    base += tot;
  }
  if (lane == 0 && !pass) counts[c] = base;
}

void launch_image_list(const Cell &cell, const CellGrid &g, double reach, const double widths[3],
                       double box_reach, const double *ax, const double *ay, const double *az,
                       int32_t *counts, const int32_t *off, float4 *img, int pass, hipStream_t st) {
  // the box test for rectangular cells (cell vectors along the axes)
  const bool rect = cell.bc == 1 && box_reach > 0.0;
  const double box2 = rect ? box_reach * box_reach : 0.0;
  const double hx = 0.5 * cell.box[AX] / g.nc[0], hy = 0.5 * cell.box[BY] / g.nc[1],
               hz = 0.5 * cell.box[CZ] / g.nc[2];
  hipLaunchKernelGGL(k_image_list, dim3((g.ncell + 3) / 4), dim3(256), 0, st, cell, g, reach * reach,
                     widths[0], widths[1], widths[2], box2, hx, hy, hz, ax, ay, az, counts, off, img,
                     pass);
}

// Neighbour lists and image lists of every cell in one walk over the atoms:
// one workgroup of four waves per cell, lane = atom, the waves taking the
// four 64-atom parts of each 256-atom tile (wave w the atoms t0 + 64 w +
// lane) and placing their entries after those of the lower waves, so each
// cell's neighbour list comes out in ascending atom order and its image list
// atom-major in that order, the images of an atom in the fixed shift order,
// exactly as k_cell_neighbours and then k_image_list over that neighbour
// list give them.  An atom is a
// neighbour when its torus distance to the cell centre is at most reach; its
// images are those k_image_list keeps (box2 > 0: within sqrt(box2) of the
// cell's box, else within reach of the centre).  pass 0 counts both lists,
// pass 1 writes them at nb_off[c] and img_off[c].  RECT (rectangular cells):
// the same distances per axis, without the zero off-diagonal products.
template <int RECT>
__global__ __launch_bounds__(256) void k_cell_lists(const Cell cell, const CellGrid g, double reach2,
                                                    double w0, double w1, double w2, double box2,
                                                    double hx, double hy, double hz,
                                                    const double *__restrict__ ax,
                                                    const double *__restrict__ ay,
                                                    const double *__restrict__ az, int n_atoms,
                                                    int32_t *__restrict__ nb_counts,
                                                    int32_t *__restrict__ img_counts,
                                                    const int32_t *__restrict__ nb_off,
                                                    int32_t *__restrict__ nb_idx,
                                                    const int32_t *__restrict__ img_off,
                                                    float4 *__restrict__ img, int pass) {
  __shared__ int s_nb[4], s_img[4];
  const int c = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long lower = (1ULL << lane) - 1ULL;
  double cx, cy, cz;
  cell_centre(cell, g, c, cx, cy, cz);
  const double reach = sqrt(reach2);
  int nb_base = pass ? nb_off[c] : 0, img_base = pass ? img_off[c] : 0;
  // this thread's atom of the next tile, loaded while the current one is worked on
  double rx = 0.0, ry = 0.0, rz = 0.0;
  if ((int)threadIdx.x < n_atoms) {
    rx = ax[threadIdx.x];
    ry = ay[threadIdx.x];
    rz = az[threadIdx.x];
  }
  for (int t0 = 0; t0 < n_atoms; t0 += 256) {
    const int j = t0 + 64 * wv + lane;
    const double dx = j < n_atoms ? rx - cx : 0.0, dy = j < n_atoms ? ry - cy : 0.0,
                 dz = j < n_atoms ? rz - cz : 0.0;
    const int jn = j + 256;
    if (jn < n_atoms) {
      rx = ax[jn];
      ry = ay[jn];
      rz = az[jn];
    }
    double s[3] = {0.0, 0.0, 0.0};
    bool in;
    if (RECT) {
      // the nearest image per axis is the torus distance's image
      s[0] = cell.inv[AX] * dx;
      s[1] = cell.inv[BY] * dy;
      s[2] = cell.inv[CZ] * dz;
      for (int d = 0; d < 3; ++d) s[d] -= rint(s[d]);
      const double x = cell.box[AX] * s[0], y = cell.box[BY] * s[1], z = cell.box[CZ] * s[2];
      in = j < n_atoms && x * x + y * y + z * z <= reach2;
    } else {
      in = j < n_atoms && torus_dist2(cell, dx, dy, dz) <= reach2;
    }
    int m = 0;
    int cand[3][2] = {{0, 0}, {0, 0}, {0, 0}}, nc[3] = {1, 1, 1};
    if (in) {
      if (!RECT) {
        s[0] = cell.inv[AX] * dx + cell.inv[BX] * dy + cell.inv[CX] * dz;
        s[1] = cell.inv[AY] * dx + cell.inv[BY] * dy + cell.inv[CY] * dz;
        s[2] = cell.inv[AZ] * dx + cell.inv[BZ] * dy + cell.inv[CZ] * dz;
        for (int d = 0; d < 3; ++d) s[d] -= rint(s[d]);
      }
      const double w[3] = {w0, w1, w2};
      for (int d = 0; d < 3; ++d)
        if ((1.0 - fabs(s[d])) * w[d] <= reach) cand[d][nc[d]++] = s[d] > 0.0 ? -1 : 1;
    }
    auto image = [&](int a, int b, int e, double &vx, double &vy, double &vz) {
      const double tx = s[0] + cand[0][a], ty = s[1] + cand[1][b], tz = s[2] + cand[2][e];
      if (RECT) {
        vx = cell.box[AX] * tx;
        vy = cell.box[BY] * ty;
        vz = cell.box[CZ] * tz;
      } else {
        vx = cell.box[AX] * tx + cell.box[BX] * ty + cell.box[CX] * tz;
        vy = cell.box[AY] * tx + cell.box[BY] * ty + cell.box[CY] * tz;
        vz = cell.box[AZ] * tx + cell.box[BZ] * ty + cell.box[CZ] * tz;
      }
      if (box2 > 0.0) {
        const double ex = fmax(fabs(vx) - hx, 0.0), ey = fmax(fabs(vy) - hy, 0.0),
                     ez = fmax(fabs(vz) - hz, 0.0);
        return ex * ex + ey * ey + ez * ez <= box2;
      }
      return vx * vx + vy * vy + vz * vz <= reach2;
    };
    if (in)
      for (int a = 0; a < nc[0]; ++a)
        for (int b = 0; b < nc[1]; ++b)
          for (int e = 0; e < nc[2]; ++e) {
            double vx, vy, vz;
            m += image(a, b, e, vx, vy, vz) ? 1 : 0;
          }
    const unsigned long long bal = __ballot(in);
    int pre = 0, tot = 0;
#pragma unroll
    for (int bit = 0; bit < 4; ++bit) {
      const unsigned long long v = __ballot((m >> bit) & 1);
      pre += __popcll(v & lower) << bit;
      tot += __popcll(v) << bit;
    }
    // entries of the lower waves' atoms of this tile come first
    int nb_pre = 0, img_pre = 0, nb_all = __popcll(bal), img_all = tot;
    if (pass) {
      if (lane == 0) {
        s_nb[wv] = __popcll(bal);
        s_img[wv] = tot;
      }
      __syncthreads();
      nb_all = img_all = 0;
      for (int v = 0; v < 4; ++v) {
        if (v < wv) {
          nb_pre += s_nb[v];
          img_pre += s_img[v];
        }
        nb_all += s_nb[v];
        img_all += s_img[v];
      }
      __syncthreads(); // the counts are read before the next tile's are written
    }
    if (pass && in) {
      nb_idx[nb_base + nb_pre + __popcll(bal & lower)] = j;
      int o = img_base + img_pre + pre;
      for (int a = 0; a < nc[0]; ++a)
        for (int b = 0; b < nc[1]; ++b)
          for (int e = 0; e < nc[2]; ++e) {
            double vx, vy, vz;
            if (image(a, b, e, vx, vy, vz))
              img[o++] = make_float4((float)vx, (float)vy, (float)vz, __int_as_float(j));
          }
    }
    nb_base += nb_all;
    img_base += img_all;
  }
  if (!pass) {
    if (lane == 0) {
      s_nb[wv] = nb_base;
      s_img[wv] = img_base;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      nb_counts[c] = s_nb[0] + s_nb[1] + s_nb[2] + s_nb[3];
      img_counts[c] = s_img[0] + s_img[1] + s_img[2] + s_img[3];
    }
  }
}

void launch_cell_lists(const Cell &cell, const CellGrid &g, double reach, const double widths[3],
                       double box_reach, const double *ax, const double *ay, const double *az, int n_atoms,
                       int32_t *nb_counts, int32_t *img_counts, const int32_t *nb_off, int32_t *nb_idx,
                       const int32_t *img_off, float4 *img, int pass, hipStream_t st) {
  const bool rect = cell.bc == 1 && box_reach > 0.0;
  const double box2 = rect ? box_reach * box_reach : 0.0;
  const double hx = 0.5 * cell.box[AX] / g.nc[0], hy = 0.5 * cell.box[BY] / g.nc[1],
               hz = 0.5 * cell.box[CZ] / g.nc[2];
  // one workgroup of four waves per cell
  if (cell.bc == 1)
    hipLaunchKernelGGL(k_cell_lists<1>, dim3(g.ncell), dim3(256), 0, st, cell, g, reach * reach,
                       widths[0], widths[1], widths[2], box2, hx, hy, hz, ax, ay, az, n_atoms, nb_counts,
                       img_counts, nb_off, nb_idx, img_off, img, pass);
  else
    hipLaunchKernelGGL(k_cell_lists<0>, dim3(g.ncell), dim3(256), 0, st, cell, g, reach * reach,
                       widths[0], widths[1], widths[2], box2, hx, hy, hz, ax, ay, az, n_atoms, nb_counts,
                       img_counts, nb_off, nb_idx, img_off, img, pass);
}

// Exclusive scans of two count arrays at once (blockIdx.x picks the pair),
// each by one workgroup as k_scan; tot[b] receives array b's total.
__global__ __launch_bounds__(1024) void k_scan2(const int32_t *__restrict__ in0, int32_t *__restrict__ off0,
                                                const int32_t *__restrict__ in1, int32_t *__restrict__ off1,
                                                int n, int32_t *__restrict__ tot) {
  __shared__ int part[1024];
  const int32_t *in = blockIdx.x ? in1 : in0;
  int32_t *off = blockIdx.x ? off1 : off0;
  const int t = threadIdx.x;
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int b = t * per, e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; ++i) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int d = 1; d < (int)blockDim.x; d <<= 1) {
    int v = (t >= d) ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t == 0) ? 0 : part[t - 1];
  for (int i = b; i < e; ++i) {
    off[i] = run;
    run += in[i];
  }
  if (t == (int)blockDim.x - 1) {
    off[n] = part[t];
    tot[blockIdx.x] = part[t];
  }
}

void launch_scan2(const int32_t *in0, int32_t *off0, const int32_t *in1, int32_t *off1, int n, int32_t *tot,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_scan2, dim3(2), dim3(1024), 0, st, in0, off0, in1, off1, n, tot);
}

// Per cell, the FP64 sum of the LJ shifts of one pseudo-atom type over the
// cell's image list: one workgroup per cell, each lane summing a strided part
// of the list in order, then a fixed pairwise tree (the sum does not depend
// on timing).
__global__ __launch_bounds__(256) void k_img_shift_sums(const CellGrid g, const AtomRec64 *__restrict__ a64,
                                                        double *__restrict__ out) {
  __shared__ double part[256];
  const int c = blockIdx.x;
  const int q0 = g.img_off[c], q1 = g.img_off[c + 1];
  double sum = 0.0;
  for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x) sum += a64[__float_as_int(g.img[q].w)].shift;
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = part[0];
}

void launch_img_shift_sums(const CellGrid &g, const AtomRec64 *a64, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_img_shift_sums, dim3(g.ncell), dim3(256), 0, st, g, a64, out);
}

__device__ __forceinline__ int cell_index(const Cell &c, const CellGrid &g, double x, double y,
                                          double z) {
  double f[3];
  f[0] = c.inv[AX] * x + c.inv[BX] * y + c.inv[CX] * z;
  f[1] = c.inv[AY] * x + c.inv[BY] * y + c.inv[CY] * z;
  f[2] = c.inv[AZ] * x + c.inv[BZ] * y + c.inv[CZ] * z;
  int ci[3];
  for (int d = 0; d < 3; ++d) {
    double u = f[d] - floor(f[d]);
    int k = (int)(u * g.nc[d]);
    ci[d] = k < 0 ? 0 : (k >= g.nc[d] ? g.nc[d] - 1 : k);
  }
  return (ci[0] * g.nc[1] + ci[1]) * g.nc[2] + ci[2];
}

// Sites binned into cell c (the counts themselves are cleared by the scan).
__device__ __forceinline__ int cell_sites(const BinWork &b, int c) { return b.start[c + 1] - b.start[c]; }

// A slot of a compacted list for each lane that calls this: one atomic per
// wave and consecutive slots for the calling lanes in lane order, so that
// the rows a wave writes by slot are contiguous (coalesced stores).
__device__ __forceinline__ int wave_slot(int32_t *counter) {
  const unsigned long long m = __ballot(1);
  const int lane = threadIdx.x & 63;
  const int leader = __builtin_ctzll(m);
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, (int)__popcll(m));
  base = __shfl(base, leader);
  return base + (int)__popcll(m & ((1ULL << lane) - 1ULL));
}

// Slots for the lanes of a workgroup that want one, with one atomic on
// counter per workgroup (a single counter shared by a whole grid serialises
// its atomics); every thread of the workgroup must call it (two barriers),
// with at most 16 waves per workgroup.  Returns -1 where want is false.
__device__ __forceinline__ int block_slot(int32_t *counter, bool want) {
  __shared__ int s_wave[16];
  __shared__ int s_base;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const unsigned long long m = __ballot(want);
  if (lane == 0) s_wave[wid] = (int)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int v = 0; v < nw; ++v) {
      const int n = s_wave[v];
      s_wave[v] = t;
      t += n;
    }
    s_base = t ? atomicAdd(counter, t) : 0;
  }
  __syncthreads();
  return want ? s_base + s_wave[wid] + (int)__popcll(m & ((1ULL << lane) - 1ULL)) : -1;
}

__global__ __launch_bounds__(256) void k_zero_i32(int32_t *__restrict__ a, int n, int32_t *__restrict__ one) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
  if (i == 0 && one) *one = 0;
}

// Lanes of a wave binned into the same cell form a group (in any lane
// order): the group's lowest lane leads it and makes one atomic for all of
// its lanes (atomics on the per-cell counters, not the arithmetic, set the
// pace of the binning).  The loop takes one distinct cell per pass, at most
// 16 passes; lanes left after that (waves of scattered sites) are groups of
// their own.
struct CellGroup {
  unsigned long long mask; // lanes of this lane's group
  int lead;                // the group's lowest lane
};

__device__ __forceinline__ CellGroup cell_group(int c, bool valid) {
  const int lane = threadIdx.x & 63;
  CellGroup r{1ULL << lane, lane};
  unsigned long long todo = __ballot(valid);
  for (int pass = 0; todo && pass < 16; ++pass) {
    const int l = __builtin_ctzll(todo);
    const int cl = __builtin_amdgcn_readlane(c, l);
    const unsigned long long m = __ballot(valid && c == cl);
    if (valid && c == cl) {
      r.mask = m;
      r.lead = l;
    }
    todo &= ~m;
  }
  return r;
}

// Counts the sites of each cell and gives each site its rank among its
// cell's sites (the value of the cell's counter its group's atomic returned
// plus its place in the group), so that the scatter needs no atomics.
// Grid-stride over the list's sites (the device-side count when the list
// has one), so that a list far below its upper bound launches no idle
// workgroups beyond the grid.
template <int BCK>
__global__ __launch_bounds__(256) void k_bin_count(const Cell cell, const CellGrid g,
                                                   const SiteList s, BinWork b) {
  const int64_t n = site_count(s);
  const int lane = threadIdx.x & 63;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n; i0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    const bool valid = i < n;
    const int64_t p = valid ? site_phys(s, i) : 0;
    double x = 0.0, y = 0.0, z = 0.0;
    if (valid) {
      x = s.x[p];
      y = s.y[p];
      z = s.z[p];
    }
    const int c = valid ? cell_index(cell, g, x, y, z) : -1;
    const CellGroup r = cell_group(c, valid);
    int base = 0;
    if (valid && lane == r.lead) base = atomicAdd(&b.count[c], (int)__popcll(r.mask));
    base = __shfl(base, r.lead);
    if (valid) {
      if (b.fixu) {
        // the offset k_bin_scatter would form, formed here from the loaded position
        double cx, cy, cz;
        cell_centre(cell, g, c, cx, cy, cz);
        double ox = x - cx, oy = y - cy, oz = z - cz;
        mi_exact<BCK>(cell, ox, oy, oz);
        b.fixu[i] = make_float4((float)ox, (float)oy, (float)oz, __int_as_float(c));
      } else {
        b.cell_of[i] = c;
      }
      b.rank[i] = base + (int)__popcll(r.mask & ((1ULL << lane) - 1ULL));
    }
  }
}

// Writes each site at start[cell] + rank: the sorted FP32 offset from its
// cell centre (image-list kernels) or its index (perm).  Grid-stride as
// k_bin_count.
template <int BCK>
__global__ __launch_bounds__(256) void k_bin_scatter(const Cell cell, const CellGrid g, const SiteList s,
                                                     BinWork b) {
  const int64_t n = site_count(s);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (b.fixu) {
      const float4 u = b.fixu[i];
      const int pos = b.start[__float_as_int(u.w)] + b.rank[i];
      // in range whenever the counts started at zero; never written outside
      if (pos < 0 || pos >= n) continue;
      b.fix[pos] = make_float4(u.x, u.y, u.z, __int_as_float((int)site_phys(s, i)));
      continue;
    }
    const int c = b.cell_of[i];
    const int pos = b.start[c] + b.rank[i];
    // in range whenever the counts started at zero; never written outside
    if (pos < 0 || pos >= n) continue;
    const int64_t site = site_phys(s, i);
    if (b.fix) {
      double cx, cy, cz;
      cell_centre(cell, g, c, cx, cy, cz);
      double ox = s.x[site] - cx, oy = s.y[site] - cy, oz = s.z[site] - cz;
      mi_exact<BCK>(cell, ox, oy, oz);
      b.fix[pos] = make_float4((float)ox, (float)oy, (float)oz, __int_as_float((int)site));
    } else {
      b.perm[pos] = (int32_t)site;
    }
  }
}

// workgroups of 256 threads for the grid-stride binning kernels: eight per
// CU of a 256-CU device at most
constexpr unsigned kBinGrid = 2048;

// exclusive scans of the per-cell site counts (start) and of the per-cell
// tile counts ceil(count/ts) (tile_off), one workgroup
// The counts are cleared once read (so that the next binning with this
// scratch starts from zero without a clearing launch), and so is the count
// of sites to recompute exactly (n_redo, nullable).
__global__ __launch_bounds__(1024) void k_scan_bins(int32_t *__restrict__ in, int n,
                                                    int32_t *__restrict__ start,
                                                    int32_t *__restrict__ tiles, int ts,
                                                    int32_t *__restrict__ n_redo) {
  // each thread a run of consecutive cells; the runs' sums scanned within
  // each wave by shuffles and across the 16 waves through LDS (two barriers)
  __shared__ int w_s[16], w_t[16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int b = t * per, e = min(n, b + per);
  int s = 0, u = 0;
  for (int i = b; i < e; ++i) {
    const int v = in[i];
    s += v;
    u += (v + ts - 1) / ts;
  }
  int is = s, iu = u; // inclusive within the wave
  for (int d = 1; d < 64; d <<= 1) {
    const int ps = __shfl_up(is, d), pu = __shfl_up(iu, d);
    if (lane >= d) {
      is += ps;
      iu += pu;
    }
  }
  if (lane == 63) {
    w_s[wid] = is;
    w_t[wid] = iu;
  }
  __syncthreads();
  if (wid == 0) {
    int vs = lane < 16 ? w_s[lane] : 0, vt = lane < 16 ? w_t[lane] : 0;
    for (int d = 1; d < 16; d <<= 1) {
      const int ps = __shfl_up(vs, d), pt = __shfl_up(vt, d);
      if (lane >= d) {
        vs += ps;
        vt += pt;
      }
    }
    if (lane < 16) {
      w_s[lane] = vs; // inclusive over the waves
      w_t[lane] = vt;
    }
  }
  __syncthreads();
  int rs = (wid ? w_s[wid - 1] : 0) + is - s, rt = (wid ? w_t[wid - 1] : 0) + iu - u;
  for (int i = b; i < e; ++i) {
    const int v = in[i];
    start[i] = rs;
    tiles[i] = rt;
    rs += v;
    rt += (v + ts - 1) / ts;
    in[i] = 0;
  }
  if (t == (int)blockDim.x - 1) {
    start[n] = w_s[15];
    tiles[n] = w_t[15];
  }
  if (t == 0 && n_redo) *n_redo = 0;
}

// k_scan_bins for at most kScanLds bins: the counts are read and cleared,
// and the starts and tile offsets written, by coalesced passes through LDS;
// each thread's run of consecutive bins is summed from LDS.
constexpr int kScanLds = 8160; // two arrays and the wave sums within 64 KiB of LDS
__global__ __launch_bounds__(1024) void k_scan_bins_lds(int32_t *__restrict__ in, int n,
                                                        int32_t *__restrict__ start,
                                                        int32_t *__restrict__ tiles, int ts,
                                                        int32_t *__restrict__ n_redo) {
  __shared__ int l_s[kScanLds], l_t[kScanLds];
  __shared__ int w_s[16], w_t[16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  for (int i = t; i < n; i += blockDim.x) {
    l_s[i] = in[i];
    in[i] = 0;
  }
  __syncthreads();
  const int per = (n + blockDim.x - 1) / blockDim.x;
  const int b = t * per, e = min(n, b + per);
  int s = 0, u = 0;
  for (int i = b; i < e; ++i) {
    const int v = l_s[i];
    s += v;
    u += (v + ts - 1) / ts;
  }
  int is = s, iu = u; // inclusive within the wave
  for (int d = 1; d < 64; d <<= 1) {
    const int ps = __shfl_up(is, d), pu = __shfl_up(iu, d);
    if (lane >= d) {
      is += ps;
      iu += pu;
    }
  }
  if (lane == 63) {
    w_s[wid] = is;
    w_t[wid] = iu;
  }
  __syncthreads();
  if (wid == 0) {
    int vs = lane < 16 ? w_s[lane] : 0, vt = lane < 16 ? w_t[lane] : 0;
    for (int d = 1; d < 16; d <<= 1) {
      const int ps = __shfl_up(vs, d), pt = __shfl_up(vt, d);
      if (lane >= d) {
        vs += ps;
        vt += pt;
      }
    }
    if (lane < 16) {
      w_s[lane] = vs;
      w_t[lane] = vt;
    }
  }
  __syncthreads();
  int rs = (wid ? w_s[wid - 1] : 0) + is - s, rt = (wid ? w_t[wid - 1] : 0) + iu - u;
  for (int i = b; i < e; ++i) {
    const int v = l_s[i];
    l_s[i] = rs;
    l_t[i] = rt;
    rs += v;
    rt += (v + ts - 1) / ts;
  }
  __syncthreads();
  for (int i = t; i < n; i += blockDim.x) {
    start[i] = l_s[i];
    tiles[i] = l_t[i];
  }
  if (t == 0) {
    start[n] = w_s[15];
    tiles[n] = w_t[15];
    if (n_redo) *n_redo = 0;
  }
}

void launch_bin_sites(const Cell &cell, const CellGrid &g, const SiteList &s, const BinWork &b,
                      hipStream_t st) {
  // counts left by an earlier binning are zero (the scan clears them)
  if (!b.counts_zero) {
    const unsigned nbc = (unsigned)((g.ncell + 255) / 256);
    hipLaunchKernelGGL(k_zero_i32, dim3(nbc), dim3(256), 0, st, b.count, g.ncell, b.n_redo);
  }
  const unsigned nb = (unsigned)std::min<int64_t>((s.n + 255) / 256, kBinGrid);
  if (nb) {
    if (cell.bc == 1)
      hipLaunchKernelGGL(k_bin_count<1>, dim3(nb), dim3(256), 0, st, cell, g, s, b);
    else if (cell.upper)
      hipLaunchKernelGGL(k_bin_count<2>, dim3(nb), dim3(256), 0, st, cell, g, s, b);
    else
      hipLaunchKernelGGL(k_bin_count<3>, dim3(nb), dim3(256), 0, st, cell, g, s, b);
  }
  if (g.ncell <= kScanLds)
    hipLaunchKernelGGL(k_scan_bins_lds, dim3(1), dim3(1024), 0, st, b.count, g.ncell, b.start, b.tile_off,
                       b.fix ? img_tile_sites(b) : 64, b.n_redo);
  else
    hipLaunchKernelGGL(k_scan_bins, dim3(1), dim3(1024), 0, st, b.count, g.ncell, b.start, b.tile_off,
                       b.fix ? img_tile_sites(b) : 64, b.n_redo);
  if (nb) {
    if (cell.bc == 1)
      hipLaunchKernelGGL(k_bin_scatter<1>, dim3(nb), dim3(256), 0, st, cell, g, s, b);
    else if (cell.upper)
      hipLaunchKernelGGL(k_bin_scatter<2>, dim3(nb), dim3(256), 0, st, cell, g, s, b);
    else
      hipLaunchKernelGGL(k_bin_scatter<3>, dim3(nb), dim3(256), 0, st, cell, g, s, b);
  }
}

// One wave per tile of <= 64 sites of one cell; the wave walks the cell's
// ascending neighbour list, so each lane adds its in-range pair terms in
// atom-index order exactly as the all-pairs loop does.
template <int PREC, int BCK, int VDW, int CC>
__global__ __launch_bounds__(256) void k_site_energy_culled(
    const Cell cell, const AtomRec64 *__restrict__ a64, const AtomRecMx *__restrict__ amx,
    const double rc2v, const double rc2c, const double alpha, const double cqa, const CellGrid g,
    const SiteList s, const BinWork b, double *__restrict__ out_v, double *__restrict__ out_c,
    uint8_t *__restrict__ out_o) {
  const int lane = threadIdx.x & 63;
  const int tile = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int ntiles = b.tile_off[g.ncell];
  if (tile >= ntiles) return;
  // wave-uniform binary search for the cell owning this tile
  int lo = 0, hi = g.ncell - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (b.tile_off[mid] <= tile) lo = mid; else hi = mid - 1;
  }
  const int c = __builtin_amdgcn_readfirstlane(lo);
  const int k = (tile - b.tile_off[c]) * 64 + lane;
  const bool live = k < cell_sites(b, c);
  const int site = live ? b.perm[b.start[c] + k] : b.perm[b.start[c]];
  const double px = s.x[site], py = s.y[site], pz = s.z[site];
  const int j0 = g.nb_off[c], j1 = g.nb_off[c + 1];
  double uv = 0.0, uc = 0.0;
  if (PREC == 0) {
#pragma clang fp contract(off)
#pragma unroll 2
    for (int q = j0; q < j1; ++q) {
      const AtomRec64 r = a64[g.nb_idx[q]];
      double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
      mi_exact<BCK>(cell, dx, dy, dz);
      double rr = dx * dx + dy * dy + dz * dz;
      if (VDW) {
        if (rr < rc2v) uv += lj_exact(r.c4, r.s2, r.shift, rr);
      }
      if (CC) {
        if (rr < rc2c) {
          double rd = sqrt(rr);
          uc += cqa * r.q * (erfc(alpha * rd) / rd);
        }
      }
    }
  } else {
    const float alphaf = (float)alpha, cqaf = (float)cqa;
#pragma unroll 4
    for (int q = j0; q < j1; ++q) {
      const AtomRecMx r = amx[g.nb_idx[q]];
      double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
      mi_fast<BCK>(cell, dx, dy, dz);
      const double rr = fma(dx, dx, fma(dy, dy, dz * dz));
      const float rrf = (float)rr;
      if (VDW) {
        const float ir = __builtin_amdgcn_rcpf(rrf);
        const float t = r.s2 * ir;
        const float r3 = t * t * t;
        const float u = fmaf(r.c4, fmaf(r3, r3, -r3), -r.shift);
        uv += (rr < rc2v) ? (double)u : 0.0;
      }
      if (CC) {
        if (rr < rc2c) {
          const float irs = __builtin_amdgcn_rsqf(rrf);
          const float rd = rrf * irs;
          const float e = erfc_mixed(alphaf * rd);
          uc += (double)((cqaf * r.q) * (e * irs));
        }
      }
    }
    const double band = 1.0e-4 * kOverlap;
    const bool near = live && ((VDW && fabs(uv - kOverlap) <= band) || (CC && fabs(uc - kOverlap) <= band));
    if (__any(near)) {
      if (near) {
#pragma clang fp contract(off)
        double ev = 0.0, ec = 0.0;
        for (int q = j0; q < j1; ++q) {
          const AtomRec64 r = a64[g.nb_idx[q]];
          double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
          mi_exact<BCK>(cell, dx, dy, dz);
          double rr = dx * dx + dy * dy + dz * dz;
          if (VDW && rr < rc2v) ev += lj_exact(r.c4, r.s2, r.shift, rr);
          if (CC && rr < rc2c) {
            double rd = sqrt(rr);
            ec += cqa * r.q * (erfc(alpha * rd) / rd);
          }
        }
        if (VDW) uv = ev;
        if (CC) uc = ec;
      }
    }
  }
  if (lane == 0 && b.pairs) {
    const int nl = min(64, cell_sites(b, c) - (tile - b.tile_off[c]) * 64);
    // counters by pair kind: 0 LJ, 1 LJ+Coulomb, 2 Coulomb
    atomicAdd(b.pairs + ((VDW && CC) ? 1 : (VDW ? 0 : 2)) * kPairLanes + (blockIdx.x & (kPairLanes - 1)),
              (unsigned long long)nl * (unsigned long long)(j1 - j0));
  }
  if (live) {
    const int64_t oi = site;
    if (out_v) out_v[oi] = uv;
    if (out_c) out_c[oi] = uc;
    if (out_o) out_o[oi] = (uint8_t)((VDW && uv >= kOverlap) || (CC && uc >= kOverlap));
  }
}

// A wave-uniform value in a VGPR (a v_mov from its SGPR): an instruction
// may read only one SGPR operand, so a value used by many instructions next
// to another SGPR operand is copied once.
__device__ __forceinline__ float vgpr_copy(float x) {
  float v;
  asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
  return v;
}

// Pair sums of NS sites per lane over a cell's image list (FP32 terms, four
// summed in FP32 and the groups in FP64).  Each wave-uniform image record and
// its parameters are loaded once (scalar loads into SGPR operands) and used
// for all NS sites of the lane: the scalar loads of the list, not the
// arithmetic, set the pace of one site per lane.
//
// VDW 1: the LJ term c12/r^12 - c6/r^6 - shift, zero beyond the cut-off.
// VDW 2 (shifted LJ only): r^2 is clamped to the cut-off instead, so that an
// entry beyond it contributes c12/rc^12 - c6/rc^6, which is its shift; the
// caller subtracts the cell's sum of shifts over the whole list once per
// site.  In-range entries then give c12/r^12 - c6/r^6 - shift as well, with
// no select and the shift out of the pair loop (the product with r^-6 folds
// into the FP32 group accumulation).
// The real-space Coulomb pair term of the image-list kernels at squared
// separation rr: CC 1 from the FP32 erfc fit (alpha rc <= 5: the fit covers
// every in-range pair; computed for all pairs and selected, like the LJ term,
// out-of-range values being finite), CC 2 from the library erfcf beyond the
// fit; q is the listed atom's charge.
template <int CC>
__device__ __forceinline__ float img_cc_term(float rr, float q, float rc2cf, float alphaf, CoulConsts ck) {
  if (CC == 1) {
    const float g = erfc_over_r(rr, __builtin_amdgcn_rsqf(rr), ck);
    return (rr < rc2cf) ? q * g : 0.0f;
  }
  float fc = 0.0f;
  if (rr < rc2cf) {
    const float irs = __builtin_amdgcn_rsqf(rr);
    const float e2 = erfc_mixed(alphaf * (rr * irs));
    fc = q * (e2 * irs);
  }
  return fc;
}

template <int NS, int VDW, int CC, int NSM>
__device__ __forceinline__ void img_sums(const float4 *__restrict__ img, const float4 *__restrict__ par,
                                         int q0, int q1, const float4 (&so)[NSM], float rc2vf,
                                         float rc2cf, float alphaf, CoulConsts ck, double cqa,
                                         double (&uv)[NSM], double (&uc)[NSM]) {
  // one pair term in FP32 (VDW 1, CC: zero beyond the cut-off; VDW 2: added
  // to fv); the LJ form keeps one SGPR operand per instruction
  auto term = [&](const float4 e, const float4 pr, const float c6v, const float4 o, float &fv, float &fc) {
    const float dx = o.x - e.x, dy = o.y - e.y, dz = o.z - e.z;
    const float rr = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (VDW == 2) {
      const float ir = __builtin_amdgcn_rcpf(fminf(rr, rc2vf));
      const float ir3 = (ir * ir) * ir;
      // c12 ir3 - c6 as one FMA: c6 is moved into a VGPR once per entry (the
      // caller's c6v), so that the FMA reads one SGPR operand
      fv = fmaf(ir3, fmaf(pr.x, ir3, -c6v), fv);
    } else if (VDW) {
      const float ir = __builtin_amdgcn_rcpf(rr);
      const float ir3 = (ir * ir) * ir;
      float y = pr.x * ir3;
      {
#pragma clang fp contract(off)
        y = y - pr.y;
      }
      const float u = fmaf(ir3, y, -pr.z);
      fv = (rr < rc2vf) ? u : 0.0f;
    }
    if (CC) fc = img_cc_term<CC>(rr, pr.w, rc2cf, alphaf, ck);
  };
  int q = q0;
  for (; q + 4 <= q1; q += 4) {
    float4 e[4], pr[4];
    float c6v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      e[u] = img[q + u];
      pr[u] = par[__float_as_int(e[u].w)];
      c6v[u] = vgpr_copy(pr[u].y);
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      float v[4] = {0.f, 0.f, 0.f, 0.f}, c[4] = {0.f, 0.f, 0.f, 0.f};
      if (VDW == 2) {
        // one FP32 accumulator through the group, in list order
#pragma unroll
        for (int u = 0; u < 4; ++u) term(e[u], pr[u], c6v[u], so[k], v[0], c[u]);
        uv[k] += (double)v[0];
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) term(e[u], pr[u], c6v[u], so[k], v[u], c[u]);
        if (VDW) uv[k] += (double)((v[0] + v[1]) + (v[2] + v[3]));
      }
      if (CC) uc[k] = fma(cqa, (double)((c[0] + c[1]) + (c[2] + c[3])), uc[k]);
    }
  }
  for (; q < q1; ++q) {
    const float4 e = img[q];
    const float4 pr = par[__float_as_int(e.w)];
    const float c6v = vgpr_copy(pr.y);
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      float v = 0.f, c = 0.f;
      term(e, pr, c6v, so[k], v, c);
      if (VDW) uv[k] += (double)v;
      if (CC) uc[k] = fma(cqa, (double)c, uc[k]);
    }
  }
}

// Mixed-precision site energies over the image lists: one wave per tile of
// <= 64 kImgLane sites of one cell (tile_off counts tiles of that size when
// the sorted offsets are written), lane l taking sites l, l + 64, ...  The site's offset
// from the cell centre and the listed image offsets are FP32 (both rounded
// once from FP64, |offset| below the cut-off plus the cell's half diagonal),
// so a pair's separation costs one subtraction per axis; the LJ term is
// c12 ir^6 - c6 ir^3 - shift and the real-space Coulomb term
// q_j erfc(alpha r)/r in FP32, the Coulomb prefactor applied to the FP64
// group sums.  Sites whose sum lands within the guard band of the overlap
// threshold are listed and recomputed exactly (FP64, atom order) by
// k_site_redo, so the overlap flags equal the FP64 path's.  CC 1: Coulomb terms from the FP32 erfc fit
// (alpha rc <= 5, branch-free); CC 2: the library erfcf beyond the fit.
// VDW 2: shifted LJ with r^2 clamped to the cut-off, cell_shift[c] the FP64
// sum of the shifts over cell c's image list (k_img_shift_sums).
template <int BCK, int VDW, int CC, int NSM>
__global__ __launch_bounds__(256) void k_site_energy_img(
    const Cell cell, const AtomRec64 *__restrict__ a64, const float4 *__restrict__ par,
    const double rc2v, const double rc2c, const double alpha, const double cqa, const CellGrid g,
    const SiteList s, const BinWork b, const double *__restrict__ cell_shift,
    double *__restrict__ out_v, double *__restrict__ out_c, uint8_t *__restrict__ out_o) {
  const int lane = threadIdx.x & 63;
  const int tile = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int ntiles = b.tile_off[g.ncell];
  if (tile >= ntiles) return;
  int lo = 0, hi = g.ncell - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (b.tile_off[mid] <= tile) lo = mid; else hi = mid - 1;
  }
  const int c = __builtin_amdgcn_readfirstlane(lo);
  const int ts = img_tile_sites(b);
  const int first = (tile - b.tile_off[c]) * ts;
  const int nl = min(ts, cell_sites(b, c) - first); // live sites of the tile
  // sorted offsets from the cell centre (written by the binning)
  float4 so[NSM];
  bool live[NSM];
#pragma unroll
  for (int k = 0; k < NSM; ++k) {
    const int kk = lane + 64 * k;
    live[k] = kk < nl;
    so[k] = b.fix[b.start[c] + first + (live[k] ? kk : 0)];
  }
  const int q0 = g.img_off[c], q1 = g.img_off[c + 1];
  double uv[NSM], uc[NSM];
#pragma unroll
  for (int k = 0; k < NSM; ++k) uv[k] = uc[k] = 0.0;
  const float rc2vf = (float)rc2v, rc2cf = (float)rc2c, alphaf = (float)alpha;
  const CoulConsts ck = coulomb_consts(alpha);
  // only as many sites per lane as the tile holds (wave-uniform)
  static_assert(NSM >= 1 && NSM <= kImgLane, "1..4 sites per lane");
  if (NSM >= 4 && nl > 192)
    img_sums<(NSM >= 4 ? 4 : 1), VDW, CC, NSM>(g.img, par, q0, q1, so, rc2vf, rc2cf, alphaf, ck, cqa, uv, uc);
  else if (NSM >= 3 && nl > 128)
    img_sums<(NSM >= 3 ? 3 : 1), VDW, CC, NSM>(g.img, par, q0, q1, so, rc2vf, rc2cf, alphaf, ck, cqa, uv, uc);
  else if (NSM >= 2 && nl > 64)
    img_sums<(NSM >= 2 ? 2 : 1), VDW, CC, NSM>(g.img, par, q0, q1, so, rc2vf, rc2cf, alphaf, ck, cqa, uv, uc);
  else
    img_sums<1, VDW, CC, NSM>(g.img, par, q0, q1, so, rc2vf, rc2cf, alphaf, ck, cqa, uv, uc);
  if (VDW == 2) {
    const double sh = cell_shift[c];
#pragma unroll
    for (int k = 0; k < NSM; ++k) uv[k] -= sh;
  }
  const double band = 1.0e-4 * kOverlap;
#pragma unroll
  for (int k = 0; k < NSM; ++k) {
    const bool near = live[k] && ((VDW && fabs(uv[k] - kOverlap) <= band) ||
                                  (CC && fabs(uc[k] - kOverlap) <= band));
    // listed for k_site_redo, which recomputes them exactly after this kernel
    if (near) b.redo[wave_slot(b.n_redo)] = make_int2(b.start[c] + first + lane + 64 * k, c);
  }
  if (lane == 0 && b.pairs)
    atomicAdd(b.pairs + ((VDW && CC) ? 1 : (VDW ? 0 : 2)) * kPairLanes + (blockIdx.x & (kPairLanes - 1)),
              (unsigned long long)nl * (unsigned long long)(q1 - q0));
#pragma unroll
  for (int k = 0; k < NSM; ++k) {
    if (live[k]) {
      const int64_t oi = __float_as_int(so[k].w);
      if (out_v) out_v[oi] = uv[k];
      if (out_c) out_c[oi] = uc[k];
      if (out_o) out_o[oi] = (uint8_t)((VDW && uv[k] >= kOverlap) || (CC && uc[k] >= kOverlap));
    }
  }
}

// Coulomb-only image-list kernel for sparse lists (the charged beads of the
// chosen orientations, some tens of sites per cell): tiles of 32 sites, two
// per wave, each half-wave walking its own cell's list, so that a cell with
// 20-40 sites fills most of 64 lanes.  The entries are vector loads (one
// address per half-wave) instead of scalar ones; each lane sums its list
// exactly as k_site_energy_img<BCK, 0, CC, 1> does (groups of four FP32 terms
// from the list start, then single terms, into the FP64 sum), so the energies
// are the same bit for bit.
template <int BCK, int CC>
__global__ __launch_bounds__(256) void k_site_energy_img_pair(
    const float4 *__restrict__ par, const double rc2c, const double alpha, const double cqa,
    const CellGrid g, const BinWork b, double *__restrict__ out_c, uint8_t *__restrict__ out_o) {
  const int lane = threadIdx.x & 63, hl = lane & 31;
  const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int ntiles = b.tile_off[g.ncell];
  if (2 * wave >= ntiles) return;
  const int tile = 2 * wave + (lane >> 5); // this half-wave's tile of 32 sites
  const bool tile_ok = tile < ntiles;
  int lo = 0, hi = g.ncell - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (b.tile_off[mid] <= min(tile, ntiles - 1)) lo = mid; else hi = mid - 1;
  }
  const int c = lo;
  const int first = (min(tile, ntiles - 1) - b.tile_off[c]) * 32;
  const int nl = tile_ok ? min(32, cell_sites(b, c) - first) : 0; // live sites of the half-tile
  const bool live = hl < nl;
  const float4 so = b.fix[b.start[c] + first + (live ? hl : 0)];
  const int q0 = g.img_off[c];
  const int len = tile_ok ? g.img_off[c + 1] - q0 : 0;
  // the longer of the two halves' lists bounds the wave's walk
  const int len_max = max(__builtin_amdgcn_readlane(len, 0), __builtin_amdgcn_readlane(len, 32));
  const float rc2cf = (float)rc2c, alphaf = (float)alpha;
  const CoulConsts ck = coulomb_consts(alpha);
  double uc = 0.0;
  for (int j = 0; j < len_max; j += 4) {
    const int m = len - j; // entries of this lane's list from j on
    if (m >= 4) {
      float4 e[4], pr[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        e[u] = g.img[q0 + j + u];
        pr[u] = par[__float_as_int(e[u].w)];
      }
      float t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float dx = so.x - e[u].x, dy = so.y - e[u].y, dz = so.z - e[u].z;
        const float rr = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        t[u] = img_cc_term<CC>(rr, pr[u].w, rc2cf, alphaf, ck);
      }
      uc = fma(cqa, (double)((t[0] + t[1]) + (t[2] + t[3])), uc);
    } else if (m > 0) {
      for (int u = 0; u < m; ++u) {
        const float4 e = g.img[q0 + j + u];
        const float4 pr = par[__float_as_int(e.w)];
        const float dx = so.x - e.x, dy = so.y - e.y, dz = so.z - e.z;
        const float rr = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        uc = fma(cqa, (double)img_cc_term<CC>(rr, pr.w, rc2cf, alphaf, ck), uc);
      }
    }
  }
  const double band = 1.0e-4 * kOverlap;
  const bool near = live && fabs(uc - kOverlap) <= band;
  // listed for k_site_redo, which recomputes them exactly after this kernel
  if (near) b.redo[wave_slot(b.n_redo)] = make_int2(b.start[c] + first + hl, c);
  if (hl == 0 && nl > 0 && b.pairs)
    atomicAdd(b.pairs + 2 * kPairLanes + (blockIdx.x & (kPairLanes - 1)),
              (unsigned long long)nl * (unsigned long long)len);
  if (live) {
    const int64_t oi = __float_as_int(so.w);
    if (out_c) out_c[oi] = uc;
    if (out_o) out_o[oi] = (uint8_t)(uc >= kOverlap);
  }
}

// Exact FP64 energies (atom order, as the culled FP64 kernel sums them) of
// the sites the image-list kernel listed as within the guard band of the
// overlap threshold: b.redo holds their sorted positions and cells.  One wave per
// listed site: the lanes evaluate 256 pair terms of the cell's ascending
// neighbour list at a time into LDS, then lane 0 adds them in list order,
// so each sum is the sequential FP64 sum bit for bit.  Keeping this rare
// path out of the image-list kernel keeps that kernel's registers (and so
// its occupancy) those of the FP32 pair loop.
template <int BCK, int VDW, int CC>
__global__ __launch_bounds__(64) void k_site_redo(const Cell cell, const AtomRec64 *__restrict__ a64,
                                                  const double rc2v, const double rc2c, const double alpha,
                                                  const double cqa, const CellGrid g, const SiteList s,
                                                  const BinWork b, double *__restrict__ out_v,
                                                  double *__restrict__ out_c, uint8_t *__restrict__ out_o) {
  constexpr int kPass = 256;
  __shared__ double tv[kPass], tc[kPass];
  const int lane = threadIdx.x;
  const int n = *b.n_redo;
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
#pragma clang fp contract(off)
    const int2 rd = b.redo[e];
    const int c = rd.y;
    const int site = __float_as_int(b.fix[rd.x].w);
    const double px = s.x[site], py = s.y[site], pz = s.z[site];
    double ev = 0.0, ec = 0.0; // lane 0's
    const int j0 = g.nb_off[c], j1 = g.nb_off[c + 1];
    for (int q0 = j0; q0 < j1; q0 += kPass) {
      // a term out of range (or past the list) is stored as +0.0: adding it
      // leaves a sum that started at +0.0 unchanged, bit for bit
      for (int u = lane; u < kPass; u += 64) {
        const int q = q0 + u;
        double v = 0.0, w = 0.0;
        if (q < j1) {
          const AtomRec64 r = a64[g.nb_idx[q]];
          double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
          mi_exact<BCK>(cell, dx, dy, dz);
          double rr = dx * dx + dy * dy + dz * dz;
          if (VDW && rr < rc2v) v = lj_exact(r.c4, r.s2, r.shift, rr);
          if (CC && rr < rc2c) {
            double rd = sqrt(rr);
            w = cqa * r.q * (erfc(alpha * rd) / rd);
          }
        }
        tv[u] = v;
        tc[u] = w;
      }
      __syncthreads();
      if (lane == 0) {
        const int m = min(kPass, j1 - q0);
        for (int u = 0; u < m; u += 8) {
          double v8[8], c8[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            v8[k] = tv[u + k];
            c8[k] = tc[u + k];
          }
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            if (VDW) ev += v8[k];
            if (CC) ec += c8[k];
          }
        }
      }
      __syncthreads();
    }
    if (lane == 0) {
      if (out_v && VDW) out_v[site] = ev;
      if (out_c && CC) out_c[site] = ec;
      if (out_o) out_o[site] = (uint8_t)((VDW && ev >= kOverlap) || (CC && ec >= kOverlap));
    }
  }
}

// A kernel launch that records ev's events around the kernel itself when
// they are given.
template <typename K, typename... Args>
static void site_dispatch(K kernel, dim3 grid, dim3 block, uint32_t lds, hipStream_t st, LaunchEvents ev,
                          Args... args) {
  if (ev.start)
    hipExtLaunchKernelGGL(kernel, grid, block, lds, st, ev.start, ev.stop, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
}

template <int BCK, int VDW, int CC>
static void img_launch(const EnergyArgs &a, const CellGrid &g, const SiteList &s, const BinWork &b,
                       const EnergyOut &o, hipStream_t st, LaunchEvents ev) {
  // upper bound of tiles: sites/(tile size) + one partial tile per cell
  const int ts = img_tile_sites(b);
  const int64_t max_tiles = s.n / ts + g.ncell + 1;
  const int w = site_waves(b);
  if constexpr (VDW == 0) {
    if (ts == 32) {
      // sparse Coulomb-only lists: two 32-site tiles per wave
      const unsigned nbp = (unsigned)(((max_tiles + 1) / 2 + w - 1) / w);
      site_dispatch(k_site_energy_img_pair<BCK, CC>, dim3(nbp), dim3(64 * w), 0, st, ev, a.par, a.rc2_cc,
                    a.alpha, a.cqa, g, b, o.u_cc, o.overlap);
      hipLaunchKernelGGL((k_site_redo<BCK, VDW, CC>), dim3(256), dim3(64), 0, st, a.cell, a.a64, a.rc2_vdw,
                         a.rc2_cc, a.alpha, a.cqa, g, s, b, o.u_vdw, o.u_cc, o.overlap);
      return;
    }
  }
  const unsigned nb = (unsigned)((max_tiles + w - 1) / w);
  // the kernel's registers are sized for the tile's sites per lane
  if (ts <= 64)
    site_dispatch(k_site_energy_img<BCK, VDW, CC, 1>, dim3(nb), dim3(64 * w), 0, st, ev, a.cell, a.a64, a.par,
                  a.rc2_vdw, a.rc2_cc, a.alpha, a.cqa, g, s, b, a.cell_shift, o.u_vdw, o.u_cc, o.overlap);
  else if (ts <= 128)
    site_dispatch(k_site_energy_img<BCK, VDW, CC, 2>, dim3(nb), dim3(64 * w), 0, st, ev, a.cell, a.a64, a.par,
                  a.rc2_vdw, a.rc2_cc, a.alpha, a.cqa, g, s, b, a.cell_shift, o.u_vdw, o.u_cc, o.overlap);
  else
    site_dispatch(k_site_energy_img<BCK, VDW, CC, 4>, dim3(nb), dim3(64 * w), 0, st, ev, a.cell, a.a64, a.par,
                  a.rc2_vdw, a.rc2_cc, a.alpha, a.cqa, g, s, b, a.cell_shift, o.u_vdw, o.u_cc, o.overlap);
  // the listed near-threshold sites, exactly (the list is short: a grid of
  // 256 waves strides over it, one wave per site)
  hipLaunchKernelGGL((k_site_redo<BCK, VDW, CC>), dim3(256), dim3(64), 0, st, a.cell, a.a64, a.rc2_vdw,
                     a.rc2_cc, a.alpha, a.cqa, g, s, b, o.u_vdw, o.u_cc, o.overlap);
}

template <int PREC, int BCK, int VDW, int CC>
static void culled_launch(const EnergyArgs &a, const CellGrid &g, const SiteList &s,
                          const BinWork &b, const EnergyOut &o, hipStream_t st, LaunchEvents ev) {
  // upper bound of tiles: sites/64 + one partial tile per cell
  const int64_t max_tiles = s.n / 64 + g.ncell + 1;
  const int w = site_waves(b);
  const unsigned nb = (unsigned)((max_tiles + w - 1) / w);
  site_dispatch(k_site_energy_culled<PREC, BCK, VDW, CC>, dim3(nb), dim3(64 * w), 0, st, ev, a.cell, a.a64,
                a.amx, a.rc2_vdw, a.rc2_cc, a.alpha, a.cqa, g, s, b, o.u_vdw, o.u_cc, o.overlap);
}

template <int VDW, int CC>
static void img_mode(const EnergyArgs &a, const CellGrid &g, const SiteList &s, const BinWork &b,
                     const EnergyOut &o, hipStream_t st, LaunchEvents ev) {
  const int bck = a.cell.bc == 1 ? 1 : (a.cell.upper ? 2 : 3);
  // CC 2: erfc arguments beyond the fit's range (alpha rc > 5) possible
  if (CC && a.alpha * sqrt(a.rc2_cc) > 5.0) {
    if (bck == 1) img_launch<1, VDW, 2>(a, g, s, b, o, st, ev);
    else if (bck == 2) img_launch<2, VDW, 2>(a, g, s, b, o, st, ev);
    else img_launch<3, VDW, 2>(a, g, s, b, o, st, ev);
  } else {
    if (bck == 1) img_launch<1, VDW, CC>(a, g, s, b, o, st, ev);
    else if (bck == 2) img_launch<2, VDW, CC>(a, g, s, b, o, st, ev);
    else img_launch<3, VDW, CC>(a, g, s, b, o, st, ev);
  }
}

template <int VDW, int CC>
static void culled_mode(const EnergyArgs &a, const CellGrid &g, const SiteList &s, const BinWork &b,
                        const EnergyOut &o, hipStream_t st, LaunchEvents ev) {
  const int bck = a.cell.bc == 1 ? 1 : (a.cell.upper ? 2 : 3);
  if (a.precision != 0 && g.img_off && a.par && b.fix) {
    // shifted LJ with the cells' shift sums: the clamped form
    if (VDW && a.cell_shift) img_mode<2, CC>(a, g, s, b, o, st, ev);
    else img_mode<VDW, CC>(a, g, s, b, o, st, ev);
    return;
  }
  if (a.precision == 0) {
    if (bck == 1) culled_launch<0, 1, VDW, CC>(a, g, s, b, o, st, ev);
    else if (bck == 2) culled_launch<0, 2, VDW, CC>(a, g, s, b, o, st, ev);
    else culled_launch<0, 3, VDW, CC>(a, g, s, b, o, st, ev);
  } else {
    if (bck == 1) culled_launch<1, 1, VDW, CC>(a, g, s, b, o, st, ev);
    else if (bck == 2) culled_launch<1, 2, VDW, CC>(a, g, s, b, o, st, ev);
    else culled_launch<1, 3, VDW, CC>(a, g, s, b, o, st, ev);
  }
}

__global__ __launch_bounds__(256) void k_deinterleave(const double *__restrict__ xyz, int64_t n,
                                                      double *__restrict__ x, double *__restrict__ y,
                                                      double *__restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = xyz[3 * i];
  y[i] = xyz[3 * i + 1];
  z[i] = xyz[3 * i + 2];
}

void launch_deinterleave(const double *xyz, int64_t n, double *x, double *y, double *z, hipStream_t st) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (nb) hipLaunchKernelGGL(k_deinterleave, dim3(nb), dim3(256), 0, st, xyz, n, x, y, z);
}

void launch_site_energy_culled(const EnergyArgs &a, const CellGrid &g, const SiteList &s,
                               const BinWork &b, const EnergyOut &o, hipStream_t st, LaunchEvents ev) {
  if (s.n <= 0 || !(a.do_vdw || a.do_cc)) {
    // nothing to launch: the events still mark this point of the stream
    if (ev.start) {
      hipEventRecord(ev.start, st);
      hipEventRecord(ev.stop, st);
    }
    return;
  }
  if (a.do_vdw && a.do_cc)
    culled_mode<1, 1>(a, g, s, b, o, st, ev);
  else if (a.do_vdw)
    culled_mode<1, 0>(a, g, s, b, o, st, ev);
  else if (a.do_cc)
    culled_mode<0, 1>(a, g, s, b, o, st, ev);
}

// ====================================================================
// Ewald: framework structure factors (src/ewald.c:1289-1731): phase power
// tables per atom, then one lane per k-vector summing over the atoms
// ====================================================================

// complex products in the reference's operation order (no contraction)
struct cx {
  double re, im;
};
__device__ __forceinline__ cx cmul(cx a, cx b) {
#pragma clang fp contract(off)
  cx r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.im * b.re + a.re * b.im;
  return r;
}
__device__ __forceinline__ cx conjc(cx a) { return cx{a.re, -a.im}; }

// Per charged atom the powers e^{i ii fx} (ii = 0..kx), e^{i jj fy}
// (jj = -ky..ky) and e^{i fz}, by the recurrences of src/ewald.c:1330-1420,
// stored [power][atom] (atom fastest).
__global__ __launch_bounds__(256) void k_sf_tables(int kx, int ky, const double *__restrict__ fx,
                                                   const double *__restrict__ fy,
                                                   const double *__restrict__ fz, int n,
                                                   double2 *__restrict__ tab) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  double s, c;
  sincos(fx[a], &s, &c);
  cx e = cx{1.0, 0.0}, e1 = cx{c, s};
  for (int i = 0; i <= kx; ++i) {
    tab[(int64_t)i * n + a] = make_double2(e.re, e.im);
    e = cmul(e, e1);
  }
  const int y0 = kx + 1 + ky; // row of jj = 0
  sincos(fy[a], &s, &c);
  e = cx{1.0, 0.0};
  e1 = cx{c, s};
  for (int j = 0; j <= ky; ++j) {
    tab[(int64_t)(y0 + j) * n + a] = make_double2(e.re, e.im);
    tab[(int64_t)(y0 - j) * n + a] = make_double2(e.re, -e.im);
    e = cmul(e, e1);
  }
  sincos(fz[a], &s, &c);
  tab[(int64_t)(kx + 2 * ky + 2) * n + a] = make_double2(c, s);
}

// One workgroup per (k_x, k_y) row of the k table: threads take atoms, form
// q e^{i(k_x x + k_y y)} from the tables and run the k_z recurrence; the
// +k_z and -k_z structure factors share four products per atom, which are
// reduced over the workgroup in a fixed order (wave shuffles, then the four
// waves in order, atom chunks in order).
constexpr int kSfMaxKz = 64;

__global__ __launch_bounds__(256) void k_sf_rows(const EwaldDev e, const double2 *__restrict__ tab,
                                                 const double *__restrict__ q, int n,
                                                 double2 *__restrict__ out) {
  __shared__ double part[4][kSfMaxKz + 1][4];
  __shared__ double acc[kSfMaxKz + 1][4];
  const int ny = 2 * e.ky + 1;
  const int row = blockIdx.x;
  const int ii = row / ny, jv = row % ny - e.ky;
  const int kzm = e.row_kz[row];
  if (kzm < 0) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < (kzm + 1) * 4; t += blockDim.x) acc[t >> 2][t & 3] = 0.0;
  __syncthreads();
  for (int a0 = 0; a0 < n; a0 += blockDim.x) {
    const int a = a0 + threadIdx.x;
    cx xy = cx{0.0, 0.0}, e1z = cx{1.0, 0.0};
    if (a < n) {
      const double2 tx = tab[(int64_t)ii * n + a];
      const double2 ty = tab[(int64_t)(e.kx + 1 + e.ky + jv) * n + a];
      const double2 tz = tab[(int64_t)(e.kx + 2 * e.ky + 2) * n + a];
      const cx r = cmul(cx{tx.x, tx.y}, cx{ty.x, ty.y});
      xy = cx{q[a] * r.re, q[a] * r.im};
      e1z = cx{tz.x, tz.y};
    }
    cx ez = cx{1.0, 0.0};
    for (int kk = 0; kk <= kzm; ++kk) {
      if (kk > 0) ez = cx{fma(ez.re, e1z.re, -ez.im * e1z.im), fma(ez.im, e1z.re, ez.re * e1z.im)};
      double v[4] = {xy.re * ez.re, xy.im * ez.im, xy.re * ez.im, xy.im * ez.re};
#pragma unroll
      for (int c = 0; c < 4; ++c)
        for (int d = 32; d > 0; d >>= 1) v[c] += __shfl_down(v[c], d);
      if (lane == 0)
        for (int c = 0; c < 4; ++c) part[wave][kk][c] = v[c];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < (kzm + 1) * 4; t += blockDim.x) {
      const int kk = t >> 2, c = t & 3;
      acc[kk][c] += ((part[0][kk][c] + part[1][kk][c]) + part[2][kk][c]) + part[3][kk][c];
    }
    __syncthreads();
  }
  // S(+k_z) = (a1 - a2, a3 + a4), S(-k_z) = (a1 + a2, a4 - a3); in the
  // origin row k = 0 is not stored and the +k_z entries sit one place lower
  const bool origin = (ii == 0 && jv == 0);
  const int off = e.row_off[row];
  for (int kk = threadIdx.x; kk <= kzm; kk += blockDim.x) {
    const double a1 = acc[kk][0], a2 = acc[kk][1], a3 = acc[kk][2], a4 = acc[kk][3];
    if (kk == 0) {
      if (!origin) out[off + kzm] = make_double2(a1 - a2, a3 + a4);
      continue;
    }
    out[off + kzm + kk - (origin ? 1 : 0)] = make_double2(a1 - a2, a3 + a4);
    out[off + kzm - kk] = make_double2(a1 + a2, a4 - a3);
  }
}

void launch_structure_factor(const EwaldDev &e, const double *fx, const double *fy, const double *fz,
                             const double *q, int n, double2 *tab, double2 *out, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_sf_tables, dim3((n + 255) / 256), dim3(256), 0, st, e.kx, e.ky, fx, fy, fz, n,
                       tab);
  const int rows = (e.kx + 1) * (2 * e.ky + 1);
  hipLaunchKernelGGL(k_sf_rows, dim3(rows), dim3(256), 0, st, e, (const double2 *)tab, q, n, out);
}



// ====================================================================
// Fourier Delta of inserted molecules (src/ewald.c:6859-7565), one lane
// per molecule.  exp(ik.r) by the reference's recurrences; the loop visits
// each (kx, |ky|, |kz|) once and forms the sign variants as conjugates,
// reading KFactor and the framework structure factor at the reference's
// k index (wave-uniform, scalar loads).
// ====================================================================

template <int ADS>
__device__ __forceinline__ double ads_term(const double2 *sads, const double4 &r, double sre, double sim) {
  const double d2 = fma(sim, sim, sre * sre);
  if (!ADS) return d2;
  const double2 st = sads[(int)r.w];
  return fma(2.0 * st.x, sre, fma(2.0 * st.y, sim, d2));
}

// Partial k-sums of one kx slice (blockIdx.y = kx) for 64 molecules per
// wave; part[kx][m] = (host, ads).  The slice's exp(i kx x) comes from the
// same recurrence as the reference (kx repeated products of exp(i x)).  The
// k-vectors are visited row by row (k_y, and for k_y > 0 the rows +k_y and
// -k_y together, which share the k_z phases), and their records {K Re S_fw,
// K Im S_fw, K, index} are read from e.kst, laid out in exactly that order
// (kst_off[kx]: the slice's first record): each k_z step reads one
// contiguous block of records through the scalar unit with one wait,
// instead of a dependent load per k-vector.
// With adsorbates present (ADS) the adsorbate term of each k is
// |S_ads + dS|^2 - |S_ads|^2 = 2 Re(S_ads* dS) + |dS|^2 (src/ewald.c:7233-7259).
template <int NC, int ADS>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) void k_fourier_part(const FourierArgs a, double2 *__restrict__ part) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.n_dev ? (int64_t)(*a.n_dev) : a.n;
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;
  const bool live = m < n;
  const int64_t mm = live ? m : 0;
  const EwaldDev &e = a.e;
  const int ii = blockIdx.y;
  constexpr int nc = NC;
  cx e1x[NC], e1y[NC], e1z[NC], ex[NC];
  double qs[NC];
#pragma unroll
  for (int t = 0; t < NC; ++t) {
    const int j = a.cidx[t];
    double px = a.px[j * a.stride + mm], py = a.py[j * a.stride + mm], pz = a.pz[j * a.stride + mm];
    qs[t] = a.q[t];
    double fx, fy, fz;
    {
#pragma clang fp contract(off)
      fx = e.cell.inv[AX] * px + e.cell.inv[BX] * py + e.cell.inv[CX] * pz;
      fy = e.cell.inv[AY] * px + e.cell.inv[BY] * py + e.cell.inv[CY] * pz;
      fz = e.cell.inv[AZ] * px + e.cell.inv[BZ] * py + e.cell.inv[CZ] * pz;
      fx *= (2.0 * M_PI);
      fy *= (2.0 * M_PI);
      fz *= (2.0 * M_PI);
    }
    double s, c;
    sincos(fx, &s, &c);
    e1x[t] = cx{c, s};
    sincos(fy, &s, &c);
    e1y[t] = cx{c, s};
    sincos(fz, &s, &c);
    e1z[t] = cx{c, s};
    ex[t] = cx{1.0, 0.0};
    for (int r = 1; r <= ii; ++r) ex[t] = (r == 1) ? e1x[t] : cmul(ex[t], e1x[t]);
    // the charge-weighted slice phase q e^{i kx x}, once per slice
    ex[t] = cx{qs[t] * ex[t].re, qs[t] * ex[t].im};
  }
  const int ny = 2 * e.ky + 1;
  {
    double host = 0.0, ads = 0.0;
    // one k-vector's contribution: host += K Re(S_fw* dS), ads += K (|dS|^2
    // or the adsorbate form)
    auto add_k = [&](const double4 &r, double sre, double sim) {
      host = fma(r.y, sim, fma(r.x, sre, host));
      ads = fma(r.z, ads_term<ADS>(a.sads, r, sre, sim), ads);
    };
    const double4 *__restrict__ p = e.kst + e.kst_off[ii];
    cx ey[NC];
#pragma unroll
    for (int t = 0; t < nc; ++t) ey[t] = cx{1.0, 0.0};
    for (int jj = 0; jj <= e.ky; ++jj) {
      if (jj > 0) {
#pragma unroll
        for (int t = 0; t < nc; ++t) ey[t] = (jj == 1) ? e1y[t] : cmul(ey[t], e1y[t]);
      }
      if (jj == 0) {
        const int row = ii * ny + e.ky;
        const int kzm = e.row_kz[row];
        if (kzm < 0) continue;
        // charge-weighted row phases q e^{i kx x} of each site (ky = 0)
        cx xy[NC], ez[NC];
#pragma unroll
        for (int t = 0; t < nc; ++t) {
          xy[t] = ex[t];
          ez[t] = cx{1.0, 0.0};
        }
        // k_z = 0 (absent in the origin row, where k = 0 is excluded)
        if (ii != 0) {
          double sre = 0.0, sim = 0.0;
#pragma unroll
          for (int t = 0; t < nc; ++t) {
            sre += xy[t].re;
            sim += xy[t].im;
          }
          add_k(p[0], sre, sim);
          ++p;
        }
        // k_z = +kk and -kk share the four products Re(xy) Re(ez), Im(xy)
        // Im(ez), Re(xy) Im(ez), Im(xy) Re(ez); records (+kk, -kk) per step
        double4 n0 = p[0], n1 = p[1];
        for (int kk = 1; kk <= kzm; ++kk) {
          const double4 r0 = n0, r1 = n1;
          p += 2;
          n0 = p[0];
          n1 = p[1];
          double a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
#pragma unroll
          for (int t = 0; t < nc; ++t) {
            ez[t] = cx{fma(ez[t].re, e1z[t].re, -ez[t].im * e1z[t].im),
                       fma(ez[t].im, e1z[t].re, ez[t].re * e1z[t].im)};
            a1 = fma(xy[t].re, ez[t].re, a1);
            a2 = fma(xy[t].im, ez[t].im, a2);
            a3 = fma(xy[t].re, ez[t].im, a3);
            a4 = fma(xy[t].im, ez[t].re, a4);
          }
          add_k(r0, a1 - a2, a3 + a4); // +k_z
          add_k(r1, a1 + a2, a4 - a3); // -k_z
        }
        continue;
      }
      // the rows +jj and -jj have the same k_z range (the cut-off depends on
      // jj^2) and share the k_z phases: four k-vectors per k_z step
      const int rp = ii * ny + (jj + e.ky);
      const int kzm = e.row_kz[rp];
      if (kzm < 0) continue;
      cx xp[NC], xm[NC], ez[NC];
#pragma unroll
      for (int t = 0; t < nc; ++t) {
        xp[t] = cmul(ex[t], ey[t]);
        xm[t] = cmul(ex[t], conjc(ey[t]));
        ez[t] = cx{1.0, 0.0};
      }
      {
        double spr = 0.0, spi = 0.0, smr = 0.0, smi = 0.0;
#pragma unroll
        for (int t = 0; t < nc; ++t) {
          spr += xp[t].re;
          spi += xp[t].im;
          smr += xm[t].re;
          smi += xm[t].im;
        }
        add_k(p[0], spr, spi);
        add_k(p[1], smr, smi);
        p += 2;
      }
      // records (+jj, +kk), (+jj, -kk), (-jj, +kk), (-jj, -kk) per step
      double4 n0 = p[0], n1 = p[1], n2 = p[2], n3 = p[3];
#pragma unroll 2
      for (int kk = 1; kk <= kzm; ++kk) {
        const double4 r0 = n0, r1 = n1, r2 = n2, r3 = n3;
        p += 4;
        n0 = p[0];
        n1 = p[1];
        n2 = p[2];
        n3 = p[3];
        double a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0, b4 = 0.0;
#pragma unroll
        for (int t = 0; t < nc; ++t) {
          ez[t] = cx{fma(ez[t].re, e1z[t].re, -ez[t].im * e1z[t].im),
                     fma(ez[t].im, e1z[t].re, ez[t].re * e1z[t].im)};
          a1 = fma(xp[t].re, ez[t].re, a1);
          a2 = fma(xp[t].im, ez[t].im, a2);
          a3 = fma(xp[t].re, ez[t].im, a3);
          a4 = fma(xp[t].im, ez[t].re, a4);
          b1 = fma(xm[t].re, ez[t].re, b1);
          b2 = fma(xm[t].im, ez[t].im, b2);
          b3 = fma(xm[t].re, ez[t].im, b3);
          b4 = fma(xm[t].im, ez[t].re, b4);
        }
        add_k(r0, a1 - a2, a3 + a4); // (+jj, +k_z)
        add_k(r1, a1 + a2, a4 - a3); // (+jj, -k_z)
        add_k(r2, b1 - b2, b3 + b4); // (-jj, +k_z)
        add_k(r3, b1 + b2, b4 - b3); // (-jj, -k_z)
      }
    }
    // the only global store, after every record load (an earlier store would
    // keep the compiler from reading the records through the scalar unit)
    if (live) part[(int64_t)ii * a.stride + m] = make_double2(host, ads);
  }
}

// Gathers the k-vector records of the Fourier kernel in its visiting order:
// kst[i] = {K S_fw.re, K S_fw.im, K, index} of k-vector idx[i], i < n.
__global__ __launch_bounds__(256) void k_kstream(const double *__restrict__ kf, const double2 *__restrict__ sf,
                                                 const int32_t *__restrict__ idx, int n,
                                                 double4 *__restrict__ kst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = idx[i];
  const double2 v = sf[k];
  const double f = kf[k];
  kst[i] = make_double4(f * v.x, f * v.y, f, (double)k);
}

void launch_kstream(const double *kf, const double2 *sf, const int32_t *idx, int n, double4 *kst,
                    hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_kstream, dim3((n + 255) / 256), dim3(256), 0, st, kf, sf, idx, n, kst);
}

// Sums the kx slices in order and adds the excluded-pair, self and
// net-charge terms (src/ewald.c:7318-7335, 7536-7565).
template <int NC>
__global__ __launch_bounds__(256) void k_fourier_final(const FourierArgs a,
                                                       const double2 *__restrict__ part) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.n_dev ? (int64_t)(*a.n_dev) : a.n;
  if (m >= n) return;
  const EwaldDev &e = a.e;
  double host = 0.0, ads = 0.0;
  for (int ii = 0; ii <= e.kx; ++ii) {
    const double2 p = part[(int64_t)ii * a.stride + m];
    host += p.x;
    ads += p.y;
  }
  double pxs[NC], pys[NC], pzs[NC], qs[NC];
  double net = 0.0;
#pragma unroll
  for (int t = 0; t < NC; ++t) {
    const int j = a.cidx[t];
    pxs[t] = a.px[j * a.stride + m];
    pys[t] = a.py[j * a.stride + m];
    pzs[t] = a.pz[j * a.stride + m];
    qs[t] = a.q[t];
    net += a.q[t];
  }
  // excluded intra pairs within each molecule: the new one's add, a removed
  // one's (charges stored negated) subtract (src/ewald.c:7318-7335, 7418-7440)
  double excl = 0.0;
#pragma unroll
  for (int p = 0; p < NC; ++p)
#pragma unroll
    for (int r = p + 1; r < NC; ++r) {
#pragma clang fp contract(off)
      if (a.grp[p] != a.grp[r]) continue;
      double dx = pxs[p] - pxs[r], dy = pys[p] - pys[r], dz = pzs[p] - pzs[r];
      double rd = sqrt(dx * dx + dy * dy + dz * dz);
      double bt0 = -erf(e.alpha * rd) / rd;
      const double t = e.C * qs[p] * qs[r] * bt0;
      excl = a.grp[p] ? excl + t : excl - t;
    }
  // net-charge terms (src/ewald.c:7536-7565)
  const double nca = a.net_q_ads;
  double dq = net - 0.0;
  double dads = ads - (excl - 0.0) - (a.self_energy - 0.0);
  dads += e.uion * ((nca + dq) * (nca + dq)) - e.uion * (nca * nca);
  double dhost = 2.0 * host;
  dhost += 2.0 * e.uion * e.net_q_fw * (nca + dq) - 2.0 * e.uion * e.net_q_fw * nca;
  a.d_host[m] = dhost;
  a.d_ads[m] = dads;
}

template <int NC>
static void fourier_nc(const FourierArgs &a, double2 *part, hipStream_t st) {
  const unsigned nb = (unsigned)((a.n + 127) / 128);
  if (a.sads)
    hipLaunchKernelGGL((k_fourier_part<NC, 1>), dim3(nb, a.e.kx + 1), dim3(128), 0, st, a, part);
  else
    hipLaunchKernelGGL((k_fourier_part<NC, 0>), dim3(nb, a.e.kx + 1), dim3(128), 0, st, a, part);
  hipLaunchKernelGGL(k_fourier_final<NC>, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a,
                     (const double2 *)part);
}

void launch_fourier(const FourierArgs &a, double2 *part, hipStream_t st) {
  if (a.n <= 0) return;
  switch (a.nc) {
  case 1: fourier_nc<1>(a, part, st); break;
  case 2: fourier_nc<2>(a, part, st); break;
  case 3: fourier_nc<3>(a, part, st); break;
  case 4: fourier_nc<4>(a, part, st); break;
  case 5: fourier_nc<5>(a, part, st); break;
  case 6: fourier_nc<6>(a, part, st); break;
  case 7: fourier_nc<7>(a, part, st); break;
  case 8: fourier_nc<8>(a, part, st); break;
  default: break;
  }
}

// ====================================================================
// Widom / CBMC
// ====================================================================
// SelectTrialPosition (src/cbmc.c:563-595) over k trials whose Boltzmann
// factor and overlap come from the callback.
template <class F>
__device__ __forceinline__ int select_trial(F bf_of, int k, double u) {
  double largest = -DBL_MAX;
  for (int t = 0; t < k; ++t) {
    double b;
    if (bf_of(t, b))
      if (b > largest) largest = b;
  }
  double sum = 0.0;
  for (int t = 0; t < k; ++t) {
    double b;
    if (bf_of(t, b)) sum += exp(b - largest);
  }
  double ws = u * sum;
  int sel = 0;
  double b0;
  double cumw = bf_of(0, b0) ? exp(b0 - largest) : 0.0;
  while (cumw < ws && sel < k - 1) {
    ++sel;
    double b;
    cumw += bf_of(sel, b) ? exp(b - largest) : 0.0;
  }
  return sel;
}

// ConvertFromABCtoXYZ (src/potentials.c:5548-5556)
__device__ __forceinline__ void abc_to_xyz(const Cell &c, double sx, double sy, double sz,
                                           double &x, double &y, double &z) {
#pragma clang fp contract(off)
  x = c.box[AX] * sx + c.box[BX] * sy + c.box[CX] * sz;
  y = c.box[AY] * sx + c.box[BY] * sy + c.box[CY] * sz;
  z = c.box[AZ] * sx + c.box[BZ] * sy + c.box[CZ] * sz;
}

// first-bead trial positions: counter draws 0..3k_fb-1 (z,y,x per trial,
// src/cbmc.c:652-667)
// BlockedPocket (src/framework.c:6246-6300) with
// ApplyBoundaryConditionUnitCell (src/potentials.c:5715-5752), FP64 in the
// reference's operation order
__device__ bool blocked_pocket(const PocketArgs &p, double x, double y, double z) {
#pragma clang fp contract(off)
  if (p.n <= 0) return false;
  for (int i = 0; i < p.n; ++i) {
    double dx = p.pk[4 * i] - x, dy = p.pk[4 * i + 1] - y, dz = p.pk[4 * i + 2] - z;
    if (p.bc == 1) {
      dx -= p.ubox[AX] * nint_d(dx / p.ubox[AX]);
      dy -= p.ubox[BY] * nint_d(dy / p.ubox[BY]);
      dz -= p.ubox[CZ] * nint_d(dz / p.ubox[CZ]);
    } else {
      const double sx = p.uinv[AX] * dx + p.uinv[BX] * dy + p.uinv[CX] * dz;
      const double sy = p.uinv[AY] * dx + p.uinv[BY] * dy + p.uinv[CY] * dz;
      const double sz = p.uinv[AZ] * dx + p.uinv[BZ] * dy + p.uinv[CZ] * dz;
      const double tx = sx - nint_d(sx), ty = sy - nint_d(sy), tz = sz - nint_d(sz);
      dx = p.ubox[AX] * tx + p.ubox[BX] * ty + p.ubox[CX] * tz;
      dy = p.ubox[AY] * tx + p.ubox[BY] * ty + p.ubox[CY] * tz;
      dz = p.ubox[AZ] * tx + p.ubox[BZ] * ty + p.ubox[CZ] * tz;
    }
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r < p.pk[4 * i + 3]) return !p.invert;
  }
  return p.invert != 0;
}

__global__ __launch_bounds__(256) void k_blocked_pockets(const PocketArgs p, const double *__restrict__ xyz,
                                                         int64_t n, uint8_t *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = blocked_pocket(p, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]) ? 1 : 0;
}


__global__ __launch_bounds__(256) void k_widom_fb_gen(const WidomArgs a, const WidomWork w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = chunk_n(a);
  if (i == 0) {
    if (a.n_eff) *a.n_eff = (int32_t)n;
    // the compaction counters of the selection kernels that follow
    if (w.n_active) *w.n_active = 0;
    if (w.n_active2) *w.n_active2 = 0;
  }
  if (i >= n) return;
  const int k = a.mol.k_fb;
  const int64_t first = chunk_first(a);
  if (a.use_host_trials) {
    for (int t = 0; t < k; ++t) {
      w.fbx[t * w.cap + i] = w.h_fb[(i * k + t) * 3 + 0];
      w.fby[t * w.cap + i] = w.h_fb[(i * k + t) * 3 + 1];
      w.fbz[t * w.cap + i] = w.h_fb[(i * k + t) * 3 + 2];
    }
    return;
  }
  const uint64_t key = stream_key(a.seed, (uint64_t)(first + i));
  uint64_t c = 0;
  for (int t = 0; t < k; ++t) {
    double sz = stream_uniform(key, c++);
    double sy = stream_uniform(key, c++);
    double sx = stream_uniform(key, c++);
    double x, y, z;
    abc_to_xyz(a.cell, sx, sy, sz, x, y, z);
    w.fbx[t * w.cap + i] = x;
    w.fby[t * w.cap + i] = y;
    w.fbz[t * w.cap + i] = z;
  }
}

// HandleFirstBead selection (src/cbmc.c:695-795) and, for molecules with
// more than one site, the orientation trials of GenerateTrialOrientations-
// SimpleSphere (src/cbmc.c:1480-1578).
// Selection of insertion i's first bead; true if it goes on (to the
// orientation trials, or, for one-site molecules, to the Fourier and
// finalize stages at position (fx, fy, fz)).
__device__ __forceinline__ bool fb_select_one(const WidomArgs &a, const WidomWork &w, int64_t i, double &fx,
                                              double &fy, double &fz) {
  const int k = a.mol.k_fb;
  const double beta = a.beta;
  uint64_t mask = 0;
  bool any = false;
  double W1 = 0.0;
  // guest-guest terms at finite loading (src/cbmc.c:683-689); zero without adsorbates
  const bool ads = w.fb_av != nullptr;
  for (int t = 0; t < k; ++t) {
    double uv = w.fb_uv[t * w.cap + i], uc = w.fb_uc[t * w.cap + i];
    const double av = ads ? w.fb_av[t * w.cap + i] : 0.0, ac = ads ? w.fb_ac[t * w.cap + i] : 0.0;
    if (uv >= kOverlap || uc >= kOverlap || av >= kOverlap || ac >= kOverlap)
      mask |= (1ULL << t);
    else {
      any = true;
      W1 += exp(-beta * (uv + av + 0.0));
    }
  }
  w.fbmask[i] = mask;
  w.ormask[i] = 0;
  w.selor[i] = 0;
  w.W2[i] = 0.0;
  w.selfb[i] = 0;
  w.W1[i] = 0.0;
  w.W[i] = 0.0;
  w.dU[i] = 0.0;
  w.uvdw[i] = 0.0;
  w.ureal[i] = 0.0;
  w.ufh[i] = 0.0;
  w.ufa[i] = 0.0;
  // a retraced configuration that itself overlaps is reported as overlap
  // (the reference would read stale trial energies there)
  if (!any || (a.retrace && (mask & 1ULL))) {
    w.stage[i] = 0;
    return false;
  }
  uint64_t key = 0;
  double usel = 0.0;
  if (a.retrace) {
    // the old position is trial 0 and is always chosen (src/cbmc.c:758-767)
  } else if (a.use_host_trials)
    usel = w.h_usel[2 * i];
  else {
    key = stream_key(a.seed, (uint64_t)(chunk_first(a) + i));
    usel = stream_uniform(key, (uint64_t)(3 * k));
  }
  auto bf = [&](int t, double &b) -> bool {
    double uv = w.fb_uv[t * w.cap + i], uc = w.fb_uc[t * w.cap + i];
    const double av = ads ? w.fb_av[t * w.cap + i] : 0.0, ac = ads ? w.fb_ac[t * w.cap + i] : 0.0;
    if (uv >= kOverlap || uc >= kOverlap || av >= kOverlap || ac >= kOverlap) return false;
    b = -beta * (uv + av + 0.0);
    return true;
  };
  int sel = a.retrace ? 0 : select_trial(bf, k, usel);
  W1 /= (double)k;
  w.selfb[i] = sel;
  w.W1[i] = W1;
  if (!a.retrace && W1 < kMinRosenbluth) {
    w.stage[i] = 1;
    return false;
  }
  fx = w.fbx[sel * w.cap + i];
  fy = w.fby[sel * w.cap + i];
  fz = w.fbz[sel * w.cap + i];
  w.selx[i] = fx;
  w.sely[i] = fy;
  w.selz[i] = fz;
  w.uvdw[i] = w.fb_uv[sel * w.cap + i];
  w.ureal[i] = w.fb_uc[sel * w.cap + i];
  if (ads) {
    w.uadsv[i] = w.fb_av[sel * w.cap + i];
    w.uadsc[i] = w.fb_ac[sel * w.cap + i];
  }
  w.W[i] = W1;
  w.stage[i] = 4;
  const MolDev &m = a.mol;
  if (m.n_sites == 1) {
    if (!a.retrace && blocked_pocket(a.pockets, fx, fy, fz)) {
      w.stage[i] = 5;
      w.W[i] = 0.0;
      w.uvdw[i] = 0.0;
      w.ureal[i] = 0.0;
      return false;
    }
  }
  return true;
}

__global__ __launch_bounds__(256) void k_widom_fb_select(const WidomArgs a, const WidomWork w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const bool on = i < chunk_n(a) && fb_select_one(a, w, i, fx, fy, fz);
  // the survivors' slots, one atomic per workgroup; orientation trials are
  // generated by k_widom_or_gen over the compacted survivors (full waves for
  // the rotations)
  const bool one = a.mol.n_sites == 1;
  const int slot = block_slot(one ? w.n_active2 : w.n_active, on);
  if (!on) return;
  if (one) {
    w.active2[slot] = (int32_t)i;
    w.cfx[slot] = fx;
    w.cfy[slot] = fy;
    w.cfz[slot] = fz;
  } else {
    w.active[slot] = (int32_t)i;
  }
}

// Orientation trials of the insertions whose first bead was selected
// (src/cbmc.c:1480-1578): one lane per active slot, the same counter
// sub-stream and draw order as if generated right after the selection.
__global__ __launch_bounds__(256) void k_widom_or_gen(const WidomArgs a, const WidomWork w) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= *w.n_active) return;
  const int64_t i = w.active[slot];
  const MolDev &m = a.mol;
  const int k = m.k_fb;
  const uint64_t key =
      (a.retrace || a.use_host_trials) ? 0 : stream_key(a.seed, (uint64_t)(chunk_first(a) + i));
  const double fx = w.selx[i], fy = w.sely[i], fz = w.selz[i];
  const int ko = m.k_or;
  uint64_t c = (uint64_t)(3 * k + 1);
  for (int iu = 0; iu < ko; ++iu) {
    Rot R;
    if (!a.use_host_trials) {
      // RandomNumberOnUnitSphere (src/utils.c:688-708); unused for rigid groups
      double r1, r2, rsq;
      int guard = 0;
      do {
        r1 = 2.0 * stream_uniform(key, c++) - 1.0;
        r2 = 2.0 * stream_uniform(key, c++) - 1.0;
        rsq = r1 * r1 + r2 * r2;
      } while (rsq >= 1.0 && ++guard < 4096);
      double X0 = stream_uniform(key, c++);
      double u1 = stream_uniform(key, c++);
      double u2 = stream_uniform(key, c++);
      R = shoemake(X0, u1, u2);
    }
    // the placed beads: body offsets from the starting bead, rotated
    for (int p = 0; p < m.n_place; ++p) {
      const int j = m.place[p];
      double cj[3];
      if (a.use_host_trials) {
        for (int d = 0; d < 3; ++d) cj[d] = w.h_cord[((i * ko + iu) * m.n_sites + j) * 3 + d];
      } else {
        double b0[3];
        b0[0] = m.body[j][0] - m.body[m.start][0];
        b0[1] = m.body[j][1] - m.body[m.start][1];
        b0[2] = m.body[j][2] - m.body[m.start][2];
        rotate(R, b0, cj);
      }
      double x, y, z;
      if (a.retrace && iu == 0) {
        // TrialPositions[0] = OldPosition (src/cbmc.c:2862-2864)
        const double *o = w.h_old + (i * m.n_sites + j) * 3;
        x = o[0];
        y = o[1];
        z = o[2];
      } else {
#pragma clang fp contract(off)
        x = fx + cj[0];
        y = fy + cj[1];
        z = fz + cj[2];
      }
      const int64_t o = ((int64_t)(m.g_start[p] + m.g_b[p]) * ko + iu) * w.cap + slot;
      w.orx[o] = x;
      w.ory[o] = y;
      w.orz[o] = z;
    }
  }
  w.usel_or[i] = a.retrace ? 0.0 : (a.use_host_trials ? w.h_usel[2 * i + 1] : stream_uniform(key, c));
}

// Rosen's selection over orientations (src/cbmc.c:1154-1478, 2762-2849)
// Rosen's selection of active slot `slot`; true if the insertion goes on
// (to the charged beads and the Fourier stage) with orientation *iwalk_out.
__device__ __forceinline__ bool or_select_one(const WidomArgs &a, const WidomWork &w, int64_t slot,
                                              int &iwalk_out) {
  const int64_t i = w.active[slot];
  const MolDev &m = a.mol;
  const int ko = m.k_or;
  const double beta = a.beta;
  const int64_t cap = w.cap;
  // guest-guest VDW of the placed beads is summed but, as in the reference,
  // not tested for overlap (src/cbmc.c:1208-1244)
  const bool ads = w.or_av != nullptr;
  auto ev_of = [&](int iu, double &ev, double &ea) -> bool {
#pragma clang fp contract(off)
    ev = 0.0;
    ea = 0.0;
    for (int p = 0; p < m.n_place; ++p) {
      // energies are stored [storage slot][orientation][active slot]
      const int64_t o = ((int64_t)(m.g_start[p] + m.g_b[p]) * ko + iu) * cap + slot;
      double uv = w.or_uv[o];
      if (uv >= kOverlap) return false;
      ev += uv;
      if (ads) ea += w.or_av[o];
    }
    return true;
  };
  uint64_t mask = 0;
  bool any = false;
  double weight = 0.0;
  // up to kOrCache orientations: each orientation's energies read once into
  // registers (fully unrolled, so the arrays stay in VGPRs) and the shifted
  // exponentials of the selection walk formed once; the same values in the
  // same order as the general path below
  constexpr int kOrCache = 16;
  double bs[kOrCache]; // -beta (U_vdw + U_ads) of each orientation
  if (ko <= kOrCache) {
#pragma unroll
    for (int iu = 0; iu < kOrCache; ++iu) {
      if (iu >= ko) break;
      double ev, ea;
      if (ev_of(iu, ev, ea)) {
        any = true;
        bs[iu] = -beta * (ev + ea + 0.0 + 0.0);
        weight += exp(bs[iu]);
      } else {
        bs[iu] = 0.0;
        mask |= (1ULL << iu);
      }
    }
  } else {
    for (int iu = 0; iu < ko; ++iu) {
      double ev, ea;
      if (ev_of(iu, ev, ea)) {
        any = true;
        weight += exp(-beta * (ev + ea + 0.0 + 0.0));
      } else
        mask |= (1ULL << iu);
    }
  }
  w.ormask[i] = mask;
  if (!any || (a.retrace && (mask & 1ULL))) {
    w.stage[i] = 2;
    w.W[i] = 0.0;
    w.uvdw[i] = 0.0;
    w.ureal[i] = 0.0;
    return false;
  }
  weight = weight / (double)ko;
  auto bf = [&](int iu, double &b) -> bool {
    double ev, ea;
    if (!ev_of(iu, ev, ea)) return false;
    b = -beta * (ev + ea + 0.0 + 0.0);
    return true;
  };
  int iwalk = 0;
  if (!a.retrace) {
    if (ko <= kOrCache) {
      // SelectTrialPosition over the cached weights (select_trial's steps);
      // bs[t] becomes exp(bs[t] - largest), 0 for an overlapping trial
      double largest = -DBL_MAX;
#pragma unroll
      for (int t = 0; t < kOrCache; ++t) {
        if (t >= ko) break;
        if (!((mask >> t) & 1ULL) && bs[t] > largest) largest = bs[t];
      }
      double sum = 0.0;
#pragma unroll
      for (int t = 0; t < kOrCache; ++t) {
        if (t >= ko) break;
        bs[t] = ((mask >> t) & 1ULL) ? 0.0 : exp(bs[t] - largest);
        if (!((mask >> t) & 1ULL)) sum += bs[t];
      }
      const double ws = w.usel_or[i] * sum;
      double cumw = bs[0];
#pragma unroll
      for (int t = 1; t < kOrCache; ++t) {
        if (t >= ko || !(cumw < ws)) break;
        iwalk = t;
        cumw += bs[t];
      }
    } else {
      iwalk = select_trial(bf, ko, w.usel_or[i]);
    }
  }
  double Wn = w.W1[i] * (weight * 1.0);
  w.selor[i] = iwalk;
  w.W2[i] = weight;
  if (!a.retrace && Wn < kMinRosenbluth) {
    w.stage[i] = 3;
    w.W[i] = 0.0;
    w.uvdw[i] = 0.0;
    w.ureal[i] = 0.0;
    return false;
  }
  if (!a.retrace && a.pockets.n > 0) {
    // every atom of the grown molecule, in atom order
    bool blocked = false;
    for (int j = 0; j < m.n_sites && !blocked; ++j) {
      double x = w.selx[i], y = w.sely[i], z = w.selz[i];
      for (int p = 0; p < m.n_place; ++p)
        if (m.place[p] == j) {
          const int64_t o = ((int64_t)(m.g_start[p] + m.g_b[p]) * ko + iwalk) * cap + slot;
          x = w.orx[o];
          y = w.ory[o];
          z = w.orz[o];
        }
      blocked = blocked_pocket(a.pockets, x, y, z);
    }
    if (blocked) {
      w.stage[i] = 5;
      w.W[i] = 0.0;
      w.uvdw[i] = 0.0;
      w.ureal[i] = 0.0;
      return false;
    }
  }
  double ev, ea;
  ev_of(iwalk, ev, ea);
  // src/cbmc.c:2819-2820
  if (ads) w.uadsv[i] = w.uadsv[i] + ea;
  w.uvdw[i] = w.uvdw[i] + ev;
  w.W[i] = Wn;
  iwalk_out = iwalk;
  return true;
}

__global__ __launch_bounds__(256) void k_widom_or_select(const WidomArgs a, const WidomWork w) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int iwalk = 0;
  const bool on = slot < *w.n_active && or_select_one(a, w, slot, iwalk);
  // slots of the charged-bead list, one atomic per workgroup
  const int s2 = block_slot(w.n_active2, on);
  if (!on) return;
  const int64_t i = w.active[slot];
  const MolDev &m = a.mol;
  const int ko = m.k_or;
  const int64_t cap = w.cap;
  w.active2[s2] = (int32_t)i;
  w.cfx[m.start * cap + s2] = w.selx[i];
  w.cfy[m.start * cap + s2] = w.sely[i];
  w.cfz[m.start * cap + s2] = w.selz[i];
  for (int p = 0; p < m.n_place; ++p) {
    const int j = m.place[p];
    const int64_t o = ((int64_t)(m.g_start[p] + m.g_b[p]) * ko + iwalk) * cap + slot;
    const double x = w.orx[o], y = w.ory[o], z = w.orz[o];
    w.cfx[j * cap + s2] = x;
    w.cfy[j * cap + s2] = y;
    w.cfz[j * cap + s2] = z;
    if (m.h_start[p] >= 0) {
      // charged placed beads of one type form one site list, rows of beads
      const int64_t h = (int64_t)(m.h_start[p] + m.h_b[p]) * cap + s2;
      w.ccx[h] = x;
      w.ccy[h] = y;
      w.ccz[h] = z;
    }
  }
}

// GrowMolecule tail and WidomAdsorbateMove bookkeeping
// (src/cbmc.c:3133-3185, src/mc_moves.c:7878-7933)
__global__ __launch_bounds__(256) void k_widom_finalize(const WidomArgs a, const WidomWork w) {
#pragma clang fp contract(off)
  const int64_t s2 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s2 >= *w.n_active2) return;
  const int64_t i = w.active2[s2];
  const MolDev &m = a.mol;
  const double beta = a.beta;
  const bool ads = w.uadsv != nullptr;
  double ucc = w.ureal[i];
  double uac = ads ? w.uadsc[i] : 0.0;
  if (m.n_sites > 1) {
    double ec = 0.0, eac = 0.0;
    for (int p = 0; p < m.n_place; ++p)
      if (m.h_start[p] >= 0) {
        const int64_t h = (int64_t)(m.h_start[p] + m.h_b[p]) * w.cap + s2;
        ec += w.cc_place[h];
        if (ads) eac += w.cc_ads[h];
      }
    ucc = ucc + ec;
    uac = uac + eac;
  }
  double W = w.W[i];
  W *= exp(-beta * (ucc + uac + 0.0));
  double fh = 0.0, fa = 0.0;
  if (a.ewald) {
    fh = w.f_host[s2];
    fa = w.f_ads[s2];
    W *= exp(-beta * (fh + fa + 0.0));
  }
  const double uv = w.uvdw[i];
  const double uav = ads ? w.uadsv[i] : 0.0;
  w.ureal[i] = ucc;
  if (ads) w.uadsc[i] = uac;
  w.ufh[i] = fh;
  w.ufa[i] = fa;
  w.W[i] = W;
  // src/mc_moves.c:7903-7913 with its zero terms left out
  w.dU[i] = uav + uv + uac + ucc + fh + fa;
}

__global__ __launch_bounds__(256) void k_widom_records(const int64_t n, const WidomWork w,
                                                       double *__restrict__ rec) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double *r = rec + i * 16;
  const int stg = w.stage[i];
  const bool acc = stg == 4;
  r[14] = (acc && w.uadsv) ? w.uadsv[i] : 0.0;
  r[15] = (acc && w.uadsc) ? w.uadsc[i] : 0.0;
  r[0] = acc ? w.W[i] : 0.0;
  r[1] = acc ? w.dU[i] : 0.0;
  r[2] = acc ? w.uvdw[i] : 0.0;
  r[3] = acc ? w.ureal[i] : 0.0;
  r[4] = acc ? w.ufh[i] : 0.0;
  r[5] = acc ? w.ufa[i] : 0.0;
  r[6] = (acc || stg == 5) ? 0.0 : 1.0; // a blocked molecule did not overlap
  r[7] = (stg >= 1) ? (double)w.selfb[i] : 0.0;
  r[8] = (stg >= 3) ? (double)w.selor[i] : 0.0;
  r[9] = (double)w.fbmask[i];
  r[10] = (double)w.ormask[i];
  r[11] = (stg >= 1) ? w.W1[i] : 0.0;
  r[12] = (stg >= 3) ? w.W2[i] : 0.0;
  r[13] = (double)stg;
}

__global__ void k_zero2(int32_t *a, int32_t *b) {
  if (threadIdx.x == 0) {
    *a = 0;
    *b = 0;
  }
}

void launch_zero_counters(int32_t *a, int32_t *b, hipStream_t st) {
  hipLaunchKernelGGL(k_zero2, dim3(1), dim3(64), 0, st, a, b);
}

static unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void launch_widom_fb_gen(const WidomArgs &a, const WidomWork &w, hipStream_t st) {
  if (a.n > 0) hipLaunchKernelGGL(k_widom_fb_gen, dim3(nblk(a.n, 256)), dim3(256), 0, st, a, w);
}
void launch_widom_fb_select(const WidomArgs &a, const WidomWork &w, hipStream_t st) {
  if (a.n > 0) hipLaunchKernelGGL(k_widom_fb_select, dim3(nblk(a.n, 256)), dim3(256), 0, st, a, w);
  if (a.n > 0 && a.mol.n_sites > 1)
    hipLaunchKernelGGL(k_widom_or_gen, dim3(nblk(a.n, 256)), dim3(256), 0, st, a, w);
}
void launch_widom_or_select(const WidomArgs &a, const WidomWork &w, hipStream_t st) {
  if (a.n > 0) hipLaunchKernelGGL(k_widom_or_select, dim3(nblk(a.n, 256)), dim3(256), 0, st, a, w);
}
void launch_widom_finalize(const WidomArgs &a, const WidomWork &w, hipStream_t st) {
  if (a.n > 0) hipLaunchKernelGGL(k_widom_finalize, dim3(nblk(a.n, 256)), dim3(256), 0, st, a, w);
}
void launch_blocked_pockets(const PocketArgs &p, const double *xyz, int64_t n, uint8_t *flags,
                            hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_blocked_pockets, dim3(nblk(n, 256)), dim3(256), 0, st, p, xyz, n, flags);
}

void launch_widom_records(const WidomArgs &a, const WidomWork &w, double *rec, hipStream_t st) {
  if (a.n > 0) hipLaunchKernelGGL(k_widom_records, dim3(nblk(a.n, 256)), dim3(256), 0, st, a.n, w, rec);
}

__global__ void k_widom_stats(const WidomArgs a, const WidomWork w, const StatArgs sa,
                              double *__restrict__ st) {
  if (threadIdx.x != 0) return;
  const double n = (double)chunk_n(a);
  const double na1 = sa.multi ? (double)*w.n_active : 0.0;
  const double na2 = (double)*w.n_active2;
  const double fb = n * sa.k_fb;
  const double orr = na1 * sa.k_or * sa.n_place;
  const double cc = sa.multi ? na2 * sa.n_cc_place : 0.0;
  st[0] += n;
  st[1] += fb;
  st[2] += orr;
  st[3] += cc;
  st[4] += (sa.ewald && sa.n_charged) ? na2 : 0.0;
  st[5] += (sa.fb_charged ? 0.0 : fb * sa.n_atoms) + orr * sa.n_atoms;
  st[6] += sa.fb_charged ? fb * sa.n_atoms : 0.0;
  st[7] += cc * sa.n_atoms;
  st[8] += (sa.ewald && sa.n_charged) ? na2 * sa.n_charged * sa.nk : 0.0;
}

void launch_widom_stats(const WidomArgs &a, const WidomWork &w, const StatArgs &sa, double *stats,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_widom_stats, dim3(1), dim3(64), 0, st, a, w, sa, stats);
}

__global__ void k_advance(int64_t *ctl, int64_t step) {
  if (threadIdx.x == 0) ctl[CTL_FIRST] += step;
}

void launch_advance(int64_t *ctl, int64_t step, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, ctl, step);
}

struct CtlVals {
  int64_t v[CTL_N];
};

__global__ void k_set_ctl(int64_t *ctl, const CtlVals c) {
  if (threadIdx.x < CTL_N) ctl[threadIdx.x] = c.v[threadIdx.x];
}

void launch_set_ctl(int64_t *ctl, int64_t first, int64_t end, const int64_t jb[6], hipStream_t st) {
  CtlVals c;
  c.v[CTL_FIRST] = first;
  c.v[CTL_END] = end;
  for (int b = 0; b < 6; ++b) c.v[CTL_JB + b] = jb[b];
  hipLaunchKernelGGL(k_set_ctl, dim3(1), dim3(64), 0, st, ctl, c);
}

// ====================================================================
// deterministic block reduction: segments of `seg` insertions, each reduced
// by one workgroup with a fixed tree; one workgroup then reduces the
// segment partials in order and adds them to the job's accumulators.
// ====================================================================
__device__ __forceinline__ int block_of(int64_t gi, const int64_t *jb) {
  int b = 0;
  while (b < 4 && gi >= jb[b + 1]) ++b;
  return b;
}

__global__ __launch_bounds__(256) void k_block_part(const WidomArgs a, const double *__restrict__ W,
                                                    const double *__restrict__ dU, int seg,
                                                    double *__restrict__ part) {
  const int64_t *jb = a.ctl + CTL_JB;
  __shared__ double sh[256][10];
  const int64_t n = chunk_n(a);
  const int64_t first = chunk_first(a);
  const int64_t s0 = (int64_t)blockIdx.x * seg;
  double acc[10];
  for (int q = 0; q < 10; ++q) acc[q] = 0.0;
  for (int64_t t = threadIdx.x; t < seg; t += blockDim.x) {
    int64_t i = s0 + t;
    if (i >= n) break;
    const int b = block_of(first + i, jb);
    double w = W[i];
    acc[2 * b] += w;
    acc[2 * b + 1] += dU[i] * w;
  }
  for (int q = 0; q < 10; ++q) sh[threadIdx.x][q] = acc[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 10; ++q) sh[threadIdx.x][q] += sh[threadIdx.x + s][q];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 10; ++q) part[(int64_t)blockIdx.x * 10 + q] = sh[0][q];
}

__global__ __launch_bounds__(256) void k_block_final(const WidomArgs a, const double *__restrict__ part,
                                                     int seg, double *__restrict__ acc15) {
  const int64_t *jb = a.ctl + CTL_JB;
  __shared__ double sh[256][10];
  const int64_t n = chunk_n(a);
  const int nseg = (int)((n + seg - 1) / seg);
  double acc[10];
  for (int q = 0; q < 10; ++q) acc[q] = 0.0;
  for (int s = threadIdx.x; s < nseg; s += blockDim.x)
    for (int q = 0; q < 10; ++q) acc[q] += part[(int64_t)s * 10 + q];
  for (int q = 0; q < 10; ++q) sh[threadIdx.x][q] = acc[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 10; ++q) sh[threadIdx.x][q] += sh[threadIdx.x + s][q];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int64_t first = chunk_first(a);
    for (int b = 0; b < 5; ++b) {
      acc15[b] += sh[0][2 * b];
      acc15[5 + b] += sh[0][2 * b + 1];
      int64_t lo = jb[b] > first ? jb[b] : first;
      int64_t hi = jb[b + 1] < first + n ? jb[b + 1] : first + n;
      acc15[10 + b] += hi > lo ? (double)(hi - lo) : 0.0;
    }
  }
}

// The end of a Widom chunk in one workgroup: the block sums of the segment
// partials (k_block_final), the chunk's work counters (k_widom_stats), then
// the control block's advance to the next chunk (k_advance), after every
// thread has read it.
__global__ __launch_bounds__(256) void k_chunk_tail(const WidomArgs a, const WidomWork w, const StatArgs sa,
                                                    const double *__restrict__ part, int seg,
                                                    double *__restrict__ acc15, double *__restrict__ stats,
                                                    int64_t *ctl, int64_t step) {
  const int64_t *jb = a.ctl + CTL_JB;
  __shared__ double sh[256][10];
  const int64_t n = chunk_n(a);
  const int64_t first = chunk_first(a);
  const int nseg = (int)((n + seg - 1) / seg);
  double acc[10];
  for (int q = 0; q < 10; ++q) acc[q] = 0.0;
  for (int s = threadIdx.x; s < nseg; s += blockDim.x)
    for (int q = 0; q < 10; ++q) acc[q] += part[(int64_t)s * 10 + q];
  for (int q = 0; q < 10; ++q) sh[threadIdx.x][q] = acc[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 10; ++q) sh[threadIdx.x][q] += sh[threadIdx.x + s][q];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  for (int b = 0; b < 5; ++b) {
    acc15[b] += sh[0][2 * b];
    acc15[5 + b] += sh[0][2 * b + 1];
    int64_t lo = jb[b] > first ? jb[b] : first;
    int64_t hi = jb[b + 1] < first + n ? jb[b + 1] : first + n;
    acc15[10 + b] += hi > lo ? (double)(hi - lo) : 0.0;
  }
  const double nd = (double)n;
  const double na1 = sa.multi ? (double)*w.n_active : 0.0;
  const double na2 = (double)*w.n_active2;
  const double fb = nd * sa.k_fb;
  const double orr = na1 * sa.k_or * sa.n_place;
  const double cc = sa.multi ? na2 * sa.n_cc_place : 0.0;
  stats[0] += nd;
  stats[1] += fb;
  stats[2] += orr;
  stats[3] += cc;
  stats[4] += (sa.ewald && sa.n_charged) ? na2 : 0.0;
  stats[5] += (sa.fb_charged ? 0.0 : fb * sa.n_atoms) + orr * sa.n_atoms;
  stats[6] += sa.fb_charged ? fb * sa.n_atoms : 0.0;
  stats[7] += cc * sa.n_atoms;
  stats[8] += (sa.ewald && sa.n_charged) ? na2 * sa.n_charged * sa.nk : 0.0;
  ctl[CTL_FIRST] += step;
}

void launch_chunk_tail(const WidomArgs &a, const WidomWork &w, const StatArgs &sa, const double *W,
                       const double *dU, double *part, int seg, double *acc, double *stats, int64_t *ctl,
                       int64_t step, hipStream_t st) {
  const int nseg = (int)((a.n + seg - 1) / seg);
  if (nseg > 0) hipLaunchKernelGGL(k_block_part, dim3(nseg), dim3(256), 0, st, a, W, dU, seg, part);
  hipLaunchKernelGGL(k_chunk_tail, dim3(1), dim3(256), 0, st, a, w, sa, (const double *)part, seg, acc, stats,
                     ctl, step);
}

void launch_block_reduce(const WidomArgs &a, const double *W, const double *dU, double *part, int seg,
                         double *acc, hipStream_t st) {
  const int nseg = (int)((a.n + seg - 1) / seg);
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_block_part, dim3(nseg), dim3(256), 0, st, a, W, dU, seg, part);
  hipLaunchKernelGGL(k_block_final, dim3(1), dim3(256), 0, st, a, (const double *)part, seg, acc);
}

// ====================================================================
// grids: MakeGrid node values (src/grids.c:530-583) from
// CalculateDerivativesAtPositionVDW (src/framework_hessian.c:63-109) and
// PotentialThirdDerivative LJ (src/potentials.c:10978-10994), reference
// operation order, FP64.
// ====================================================================
template <int BCK>
__device__ __forceinline__ void derivs_at(const GridArgs &g, double px, double py, double pz,
                                          double o[14]) {
#pragma clang fp contract(off)
  double value = 0, fx = 0, fy = 0, fz = 0, hax = 0, hbx = 0, hcx = 0, hay = 0, hby = 0, hcy = 0,
         haz = 0, hbz = 0, hcz = 0, third = 0;
  for (int j = 0; j < g.n_atoms; ++j) {
    const AtomRec64 r = g.a64[j];
    double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    mi_exact<BCK>(g.cell, dx, dy, dz);
    double rr = dx * dx + dy * dy + dz * dz;
    if (rr < g.rc2) {
      double t = r.s2 / rr;
      double rri3 = (t * t) * t;
      double F = r.c4 * (rri3 * (rri3 - 1.0)) - r.shift;
      double DF = g.c24[j] * (rri3 * (1.0 - 2.0 * rri3)) / rr;
      double DDF = g.c96[j] * (rri3 * (7.0 * rri3 - 2.0)) / (rr * rr);
      double DDDF = g.c384[j] * (rri3 * (5.0 - 28.0 * rri3)) / ((rr * rr) * (rr * rr));
      value += F;
      fx += dx * DF;
      fy += dy * DF;
      fz += dz * DF;
      hax += DDF * dx * dx + DF;
      hbx += DDF * dy * dx;
      hcx += DDF * dz * dx;
      hay += DDF * dx * dy;
      hby += DDF * dy * dy + DF;
      hcy += DDF * dz * dy;
      haz += DDF * dx * dz;
      hbz += DDF * dy * dz;
      hcz += DDF * dz * dz + DF;
      third += DDDF * dx * dy * dz;
    }
  }
  o[0] = value;
  o[1] = fx;
  o[2] = fy;
  o[3] = fz;
  o[4] = hax;
  o[5] = hay;
  o[6] = haz;
  o[7] = hbx;
  o[8] = hby;
  o[9] = hbz;
  o[10] = hcx;
  o[11] = hcy;
  o[12] = hcz;
  o[13] = third;
}

// MakeGrid's node record (src/grids.c:520-565, 652-676): the cap (above
// the overlap energy; for the Coulomb grid also an atom closer than 1 A),
// then the 8 channels scaled to the unit cube, FP32 (RASPA2's layout) or FP64
__device__ __forceinline__ void store_node(const GridArgs &g, int64_t node, double value, double f0,
                                           double f1, double f2, double hay, double haz, double hbz,
                                           double third, bool cap) {
#pragma clang fp contract(off)
  if (cap) {
    value = 2.0 * kOverlap;
    f0 = fmin(fmax(f0, -kOverlap), kOverlap);
    f1 = fmin(fmax(f1, -kOverlap), kOverlap);
    f2 = fmin(fmax(f2, -kOverlap), kOverlap);
    hay = haz = hbz = 0.0;
    third = 0.0;
  }
  const double dx = g.delta[0], dy = g.delta[1], dz = g.delta[2];
  if (g.fp64) {
    double *o = (double *)g.out + node * 8;
    o[0] = value;
    o[1] = f0 * dx;
    o[2] = f1 * dy;
    o[3] = f2 * dz;
    o[4] = hay * (dx * dy);
    o[5] = haz * (dx * dz);
    o[6] = hbz * (dy * dz);
    o[7] = third * (dx * dy * dz);
  } else {
    float c[8];
    c[0] = (float)value;
    c[1] = (float)((double)(float)f0 * dx);
    c[2] = (float)((double)(float)f1 * dy);
    c[3] = (float)((double)(float)f2 * dz);
    c[4] = (float)((double)(float)hay * (dx * dy));
    c[5] = (float)((double)(float)haz * (dx * dz));
    c[6] = (float)((double)(float)hbz * (dy * dz));
    c[7] = (float)((double)(float)third * (dx * dy * dz));
    float4 *o = (float4 *)((float *)g.out + node * 8);
    o[0] = make_float4(c[0], c[1], c[2], c[3]);
    o[1] = make_float4(c[4], c[5], c[6], c[7]);
  }
}

__device__ __forceinline__ void node_position(const GridArgs &g, int64_t node, double &px, double &py,
                                              double &pz) {
#pragma clang fp contract(off)
  const int64_t nz = g.n[2] + 1, ny = g.n[1] + 1;
  const int k = (int)(node % nz);
  const int j = (int)((node / nz) % ny);
  const int i = (int)(node / (nz * ny));
  px = i * g.size[0] / g.n[0] + g.shift[0];
  py = j * g.size[1] / g.n[1] + g.shift[1];
  pz = k * g.size[2] / g.n[2] + g.shift[2];
}

template <int BCK>
__global__ __launch_bounds__(256) void k_grid_build(const GridArgs g) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n_nodes) return;
  const int64_t node = g.node0 + t;
  double px, py, pz;
  node_position(g, node, px, py, pz);
  double d[14];
  derivs_at<BCK>(g, px, py, pz, d);
  store_node(g, node, d[0], d[1], d[2], d[3], d[5], d[6], d[9], d[13], d[0] > kOverlap);
}

// CalculateDerivativesAtPositionReal (src/framework_hessian.c:111-175): the
// real-space Ewald pair term of a unit charge and its first three radial
// derivative factors, FP64 in the reference's operation order (the probe
// charge 1.0 drops out of every product exactly)
__device__ __forceinline__ void real_terms(const GridArgs &g, double qb, double rr, double &r,
                                           double &F, double &DF, double &DDF, double &DDDF) {
#pragma clang fp contract(off)
  const double alpha = g.alpha, C = g.cc, sp = g.sqrtpi;
  r = sqrt(rr);
  const double ec = erfc(alpha * r);
  const double a2 = alpha * alpha;
  const double ex = exp(-a2 * rr);
  F = C * qb * (ec / r);
  DF = -C * qb * ((2.0 * alpha) * r * ex / sp + ec) / (r * rr);
  DDF = qb * C * ((2.0 * alpha) * ex * r * (3.0 + (2.0 * a2) * rr) / sp + 3.0 * ec) / ((rr * rr) * r);
  const double ar = a2 * rr;
  DDDF = qb * C *
         ((-2.0 * alpha) * ex * r * ((15.0 + (10.0 * a2) * rr) + 4.0 * (ar * ar)) / sp - 15.0 * ec) /
         (((rr * rr) * rr) * r);
}

// the grid's 8 channels of the real-space potential over atoms idx[j0..j1)
// (idx null: atoms j0..j1), in that order; returns the smallest in-range
// distance (DBL_MAX when none)
template <int BCK>
__device__ __forceinline__ double coulomb_channels(const GridArgs &g, const int32_t *idx, int j0,
                                                   int j1, double px, double py, double pz,
                                                   double o[8]) {
#pragma clang fp contract(off)
  double value = 0, fx = 0, fy = 0, fz = 0, hay = 0, haz = 0, hbz = 0, third = 0;
  double smallest = DBL_MAX;
  for (int q = j0; q < j1; ++q) {
    const int j = idx ? idx[q] : q;
    const double4 a = g.aq[j];
    double dx = px - a.x, dy = py - a.y, dz = pz - a.z;
    mi_exact<BCK>(g.cell, dx, dy, dz);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < g.rc2) {
      double r, F, DF, DDF, DDDF;
      real_terms(g, a.w, rr, r, F, DF, DDF, DDDF);
      if (r < smallest) smallest = r;
      value += F;
      fx += dx * DF;
      fy += dy * DF;
      fz += dz * DF;
      hay += DDF * dx * dy;
      haz += DDF * dx * dz;
      hbz += DDF * dy * dz;
      third += DDDF * dx * dy * dz;
    }
  }
  o[0] = value;
  o[1] = fx;
  o[2] = fy;
  o[3] = fz;
  o[4] = hay;
  o[5] = haz;
  o[6] = hbz;
  o[7] = third;
  return smallest;
}

template <int BCK>
__global__ __launch_bounds__(256) void k_coulomb_grid_build(const GridArgs g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n_nodes) return;
  const int64_t node = g.node0 + t;
  double px, py, pz, o[8];
  node_position(g, node, px, py, pz);
  const double rmin = coulomb_channels<BCK>(g, nullptr, 0, g.n_atoms, px, py, pz, o);
  store_node(g, node, o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], o[0] > kOverlap || rmin < 1.0);
}

template <int BCK>
__global__ __launch_bounds__(256) void k_derivs_real(const GridArgs g, const double *__restrict__ xyz,
                                                     int64_t n, double *__restrict__ out,
                                                     double *__restrict__ smallest) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double px = xyz[3 * t], py = xyz[3 * t + 1], pz = xyz[3 * t + 2];
  double value = 0, fx = 0, fy = 0, fz = 0, h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, third = 0;
  double rmin = DBL_MAX;
  for (int j = 0; j < g.n_atoms; ++j) {
    const double4 a = g.aq[j];
    double dx = px - a.x, dy = py - a.y, dz = pz - a.z;
    mi_exact<BCK>(g.cell, dx, dy, dz);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < g.rc2) {
      double r, F, DF, DDF, DDDF;
      real_terms(g, a.w, rr, r, F, DF, DDF, DDDF);
      if (r < rmin) rmin = r;
      value += F;
      fx += dx * DF;
      fy += dy * DF;
      fz += dz * DF;
      h[AX] += DDF * dx * dx + DF;
      h[BX] += DDF * dy * dx;
      h[CX] += DDF * dz * dx;
      h[AY] += DDF * dx * dy;
      h[BY] += DDF * dy * dy + DF;
      h[CY] += DDF * dz * dy;
      h[AZ] += DDF * dx * dz;
      h[BZ] += DDF * dy * dz;
      h[CZ] += DDF * dz * dz + DF;
      third += DDDF * dx * dy * dz;
    }
  }
  double *o = out + t * 14;
  o[0] = value;
  o[1] = fx;
  o[2] = fy;
  o[3] = fz;
  for (int q = 0; q < 9; ++q) o[4 + q] = h[q];
  o[13] = third;
  smallest[t] = rmin;
}

void launch_derivatives_real(const GridArgs &g, const double *xyz, int64_t n, double *out,
                             double *smallest, hipStream_t st) {
  const unsigned nb = nblk(n, 256);
  if (!nb) return;
  if (g.cell.bc == 1)
    hipLaunchKernelGGL(k_derivs_real<1>, dim3(nb), dim3(256), 0, st, g, xyz, n, out, smallest);
  else if (g.cell.upper)
    hipLaunchKernelGGL(k_derivs_real<2>, dim3(nb), dim3(256), 0, st, g, xyz, n, out, smallest);
  else
    hipLaunchKernelGGL(k_derivs_real<3>, dim3(nb), dim3(256), 0, st, g, xyz, n, out, smallest);
}

// one thread per node: 32 contiguous bytes on the node-major side, eight
// coalesced 4-byte streams on the channel-major side
__global__ __launch_bounds__(256) void k_grid_transpose(const float *__restrict__ src,
                                                        float *__restrict__ dst, int64_t nodes,
                                                        int to_file) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nodes) return;
  if (to_file) {
    const float4 *p = (const float4 *)(src + t * 8);
    const float4 a = p[0], b = p[1];
    dst[0 * nodes + t] = a.x;
    dst[1 * nodes + t] = a.y;
    dst[2 * nodes + t] = a.z;
    dst[3 * nodes + t] = a.w;
    dst[4 * nodes + t] = b.x;
    dst[5 * nodes + t] = b.y;
    dst[6 * nodes + t] = b.z;
    dst[7 * nodes + t] = b.w;
  } else {
    float4 *p = (float4 *)(dst + t * 8);
    p[0] = make_float4(src[0 * nodes + t], src[1 * nodes + t], src[2 * nodes + t], src[3 * nodes + t]);
    p[1] = make_float4(src[4 * nodes + t], src[5 * nodes + t], src[6 * nodes + t], src[7 * nodes + t]);
  }
}

void launch_grid_transpose(const float *src, float *dst, int64_t nodes, int to_file, hipStream_t st) {
  const unsigned nb = nblk(nodes, 256);
  if (nb) hipLaunchKernelGGL(k_grid_transpose, dim3(nb), dim3(256), 0, st, src, dst, nodes, to_file);
}

template <int BCK>
__global__ __launch_bounds__(256) void k_derivs(const GridArgs g, const double *__restrict__ xyz,
                                                int64_t n, double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;
  const int64_t m = t < n ? t : 0;
  double d[14];
  derivs_at<BCK>(g, xyz[3 * m], xyz[3 * m + 1], xyz[3 * m + 2], d);
  if (t < n)
    for (int q = 0; q < 14; ++q) out[t * 14 + q] = d[q];
}

__global__ __launch_bounds__(256) void k_grid_nodes(const GridArgs g, double *__restrict__ x,
                                                    double *__restrict__ y, double *__restrict__ z) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.n_nodes) return;
  double px, py, pz;
  node_position(g, g.node0 + t, px, py, pz);
  x[t] = px;
  y[t] = py;
  z[t] = pz;
}

void launch_grid_nodes(const GridArgs &g, double *x, double *y, double *z, hipStream_t st) {
  const unsigned nb = nblk(g.n_nodes, 256);
  if (nb) hipLaunchKernelGGL(k_grid_nodes, dim3(nb), dim3(256), 0, st, g, x, y, z);
}

// derivative channels of the grid (value, dU/dx, dU/dy, dU/dz, d2U/dxdy,
// d2U/dxdz, d2U/dydz, d3U/dxdydz) over one cell's neighbour list, in atom
// order, FP64 with the reference's operations (as derivs_at)
template <int BCK>
__device__ __forceinline__ void node_channels_exact(const GridArgs &g, const CellGrid &cg, int c,
                                                    double px, double py, double pz, double o[8]) {
#pragma clang fp contract(off)
  double value = 0, fx = 0, fy = 0, fz = 0, hay = 0, haz = 0, hbz = 0, third = 0;
  const int j0 = cg.nb_off[c], j1 = cg.nb_off[c + 1];
  for (int q = j0; q < j1; ++q) {
    const int j = cg.nb_idx[q];
    const AtomRec64 r = g.a64[j];
    double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    mi_exact<BCK>(g.cell, dx, dy, dz);
    double rr = dx * dx + dy * dy + dz * dz;
    if (rr < g.rc2) {
      double t = r.s2 / rr;
      double rri3 = (t * t) * t;
      double F = r.c4 * (rri3 * (rri3 - 1.0)) - r.shift;
      double DF = g.c24[j] * (rri3 * (1.0 - 2.0 * rri3)) / rr;
      double DDF = g.c96[j] * (rri3 * (7.0 * rri3 - 2.0)) / (rr * rr);
      double DDDF = g.c384[j] * (rri3 * (5.0 - 28.0 * rri3)) / ((rr * rr) * (rr * rr));
      value += F;
      fx += dx * DF;
      fy += dy * DF;
      fz += dz * DF;
      hay += DDF * dx * dy;
      haz += DDF * dx * dz;
      hbz += DDF * dy * dz;
      third += DDDF * dx * dy * dz;
    }
  }
  o[0] = value;
  o[1] = fx;
  o[2] = fy;
  o[3] = fz;
  o[4] = hay;
  o[5] = haz;
  o[6] = hbz;
  o[7] = third;
}

template <int PREC, int BCK>
__global__ __launch_bounds__(256) void k_grid_culled(const GridArgs g, const CellGrid cg,
                                                     const SiteList s, const BinWork b,
                                                     const float4 *__restrict__ par) {
  const int lane = threadIdx.x & 63;
  const int tile = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int ntiles = b.tile_off[cg.ncell];
  if (tile >= ntiles) return;
  int lo = 0, hi = cg.ncell - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (b.tile_off[mid] <= tile) lo = mid; else hi = mid - 1;
  }
  const int c = __builtin_amdgcn_readfirstlane(lo);
  const int k = (tile - b.tile_off[c]) * 64 + lane;
  const bool live = k < cell_sites(b, c);
  const int site = live ? b.perm[b.start[c] + k] : b.perm[b.start[c]];
  const double px = s.x[site], py = s.y[site], pz = s.z[site];
  double o[8];
  bool cap;
  if (PREC == 2) {
    const double rmin = coulomb_channels<BCK>(g, cg.nb_idx, cg.nb_off[c], cg.nb_off[c + 1], px, py, pz, o);
    cap = o[0] > kOverlap || rmin < 1.0;
  } else if (PREC == 0) {
    node_channels_exact<BCK>(g, cg, c, px, py, pz, o);
    cap = o[0] > kOverlap;
  } else {
    double cx, cy, cz;
    cell_centre(g.cell, cg, c, cx, cy, cz);
    double ox = px - cx, oy = py - cy, oz = pz - cz;
    mi_exact<BCK>(g.cell, ox, oy, oz);
    const float fx = (float)ox, fy = (float)oy, fz = (float)oz;
    const float rc2f = (float)g.rc2;
    for (int q = 0; q < 8; ++q) o[q] = 0.0;
    const int q0 = cg.img_off[c], q1 = cg.img_off[c + 1];
    for (int q = q0; q < q1; q += 4) {
      float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (q + u < q1) {
          const float4 e = cg.img[q + u];
          const float4 pr = par[__float_as_int(e.w)];
          const float dx = fx - e.x, dy = fy - e.y, dz = fz - e.z;
          const float rr = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          if (rr < rc2f) {
            const float ir = __builtin_amdgcn_rcpf(rr);
            const float ir2 = ir * ir;
            const float ir3 = ir2 * ir;
            // with s = (sigma/r)^6: A = 4 eps s^2 = c12/r^12, B = 4 eps s = c6/r^6,
            // U = A - B - shift and the derivative factors
            // 24 eps s (1 - 2 s)/r^2, 96 eps s (7 s - 2)/r^4, 384 eps s (5 - 28 s)/r^6
            const float B = pr.y * ir3;
            const float A = (pr.x * ir3) * ir3;
            const float F = (A - B) - pr.z;
            const float DF = fmaf(-12.0f, A, 6.0f * B) * ir;
            const float DDF = fmaf(168.0f, A, -48.0f * B) * ir2;
            const float DDDF = fmaf(-2688.0f, A, 480.0f * B) * (ir2 * ir2);
            const float ddx = DDF * dx;
            a[0] += F;
            a[1] = fmaf(dx, DF, a[1]);
            a[2] = fmaf(dy, DF, a[2]);
            a[3] = fmaf(dz, DF, a[3]);
            a[4] = fmaf(ddx, dy, a[4]);
            a[5] = fmaf(ddx, dz, a[5]);
            a[6] = fmaf(DDF * dy, dz, a[6]);
            a[7] = fmaf(DDDF * dx * dy, dz, a[7]);
          }
        }
      }
#pragma unroll
      for (int ch = 0; ch < 8; ++ch) o[ch] += (double)a[ch];
    }
    // nodes within the guard band of the cap are recomputed exactly
    const bool near = live && fabs(o[0] - kOverlap) <= 1.0e-4 * kOverlap;
    if (__any(near))
      if (near) node_channels_exact<BCK>(g, cg, c, px, py, pz, o);
    cap = o[0] > kOverlap;
  }
  if (lane == 0 && b.pairs) {
    // executed (node, list entry) evaluations: image entries in the mixed
    // form, neighbour atoms otherwise
    const int nl = min(64, cell_sites(b, c) - (tile - b.tile_off[c]) * 64);
    const int len = PREC == 1 ? cg.img_off[c + 1] - cg.img_off[c] : cg.nb_off[c + 1] - cg.nb_off[c];
    atomicAdd(b.pairs + (PREC == 2 ? 1 : 0) * kPairLanes + (blockIdx.x & (kPairLanes - 1)),
              (unsigned long long)nl * (unsigned long long)len);
  }
  if (live) store_node(g, g.node0 + site, o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7], cap);
}

void launch_grid_culled(const GridArgs &g, const CellGrid &cg, const SiteList &s, const BinWork &b,
                        const float4 *par, int precision, hipStream_t st) {
  const int64_t max_tiles = s.n / 64 + cg.ncell + 1;
  const unsigned nb = (unsigned)((max_tiles + 3) / 4);
  const int bck = g.cell.bc == 1 ? 1 : (g.cell.upper ? 2 : 3);
  if (g.coulomb) {
    if (bck == 1) hipLaunchKernelGGL((k_grid_culled<2, 1>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
    else if (bck == 2) hipLaunchKernelGGL((k_grid_culled<2, 2>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
    else hipLaunchKernelGGL((k_grid_culled<2, 3>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
  } else if (precision == 0) {
    if (bck == 1) hipLaunchKernelGGL((k_grid_culled<0, 1>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
    else if (bck == 2) hipLaunchKernelGGL((k_grid_culled<0, 2>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
    else hipLaunchKernelGGL((k_grid_culled<0, 3>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
  } else {
    if (bck == 1) hipLaunchKernelGGL((k_grid_culled<1, 1>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
    else if (bck == 2) hipLaunchKernelGGL((k_grid_culled<1, 2>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
    else hipLaunchKernelGGL((k_grid_culled<1, 3>), dim3(nb), dim3(256), 0, st, g, cg, s, b, par);
  }
}

void launch_grid_build(const GridArgs &g, hipStream_t st) {
  unsigned nb = nblk(g.n_nodes, 256);
  if (!nb) return;
  if (g.coulomb) {
    if (g.cell.bc == 1)
      hipLaunchKernelGGL(k_coulomb_grid_build<1>, dim3(nb), dim3(256), 0, st, g);
    else if (g.cell.upper)
      hipLaunchKernelGGL(k_coulomb_grid_build<2>, dim3(nb), dim3(256), 0, st, g);
    else
      hipLaunchKernelGGL(k_coulomb_grid_build<3>, dim3(nb), dim3(256), 0, st, g);
    return;
  }
  if (g.cell.bc == 1)
    hipLaunchKernelGGL(k_grid_build<1>, dim3(nb), dim3(256), 0, st, g);
  else if (g.cell.upper)
    hipLaunchKernelGGL(k_grid_build<2>, dim3(nb), dim3(256), 0, st, g);
  else
    hipLaunchKernelGGL(k_grid_build<3>, dim3(nb), dim3(256), 0, st, g);
}

void launch_derivatives(const GridArgs &g, const double *xyz, int64_t n, double *out,
                        hipStream_t st) {
  unsigned nb = nblk(n, 256);
  if (!nb) return;
  if (g.cell.bc == 1)
    hipLaunchKernelGGL(k_derivs<1>, dim3(nb), dim3(256), 0, st, g, xyz, n, out);
  else if (g.cell.upper)
    hipLaunchKernelGGL(k_derivs<2>, dim3(nb), dim3(256), 0, st, g, xyz, n, out);
  else
    hipLaunchKernelGGL(k_derivs<3>, dim3(nb), dim3(256), 0, st, g, xyz, n, out);
}

// ====================================================================
// InterpolateVDWGrid (src/grids.c:979-1077).  The Lekien-Marsden tricubic
// with corner data (f, fx, fy, fz, fxy, fxz, fyz, fxyz) equals the tensor
// product of 1D cubic Hermite bases, evaluated here directly (64 products)
// instead of the 64x64 matrix-vector product.
// ====================================================================
__device__ __forceinline__ void hermite(double t, double h[2][2]) {
  double t2 = t * t, t3 = t2 * t;
  h[0][0] = 2.0 * t3 - 3.0 * t2 + 1.0; // value at corner 0
  h[0][1] = t3 - 2.0 * t2 + t;         // derivative at corner 0
  h[1][0] = -2.0 * t3 + 3.0 * t2;      // value at corner 1
  h[1][1] = t3 - t2;                   // derivative at corner 1
}

__device__ double interp_at(const InterpArgs &a, double px, double py, double pz) {
  double pos[3] = {px, py, pz};
  {
#pragma clang fp contract(off)
    if (a.bc == 1) {
      for (int d = 0; d < 3; ++d) {
        pos[d] -= a.ulen[d] * nint_d(pos[d] / a.ulen[d]);
        if (pos[d] < 0.0) pos[d] += a.ulen[d];
      }
    } else {
      double s[3];
      s[0] = a.uinv[AX] * pos[0] + a.uinv[BX] * pos[1] + a.uinv[CX] * pos[2];
      s[1] = a.uinv[AY] * pos[0] + a.uinv[BY] * pos[1] + a.uinv[CY] * pos[2];
      s[2] = a.uinv[AZ] * pos[0] + a.uinv[BZ] * pos[1] + a.uinv[CZ] * pos[2];
      for (int d = 0; d < 3; ++d) {
        s[d] -= nint_d(s[d]);
        if (s[d] < 0.0) s[d] += 1.0;
      }
      pos[0] = a.ubox[AX] * s[0] + a.ubox[BX] * s[1] + a.ubox[CX] * s[2];
      pos[1] = a.ubox[AY] * s[0] + a.ubox[BY] * s[1] + a.ubox[CY] * s[2];
      pos[2] = a.ubox[AZ] * s[0] + a.ubox[BZ] * s[1] + a.ubox[CZ] * s[2];
    }
  }
  double sv[3];
  int c0[3];
  double r[3];
  for (int d = 0; d < 3; ++d) {
    sv[d] = ((pos[d] - a.shift[d]) / a.size[d]) * (double)a.n[d];
    c0[d] = (int)sv[d];
    if (c0[d] < 0) c0[d] = 0;
    if (c0[d] > a.n[d] - 1) c0[d] = a.n[d] - 1;
    r[d] = sv[d] - c0[d];
  }
  const int64_t ny = a.n[1] + 1, nz = a.n[2] + 1;
  double X[8][8]; // [corner][channel]
  for (int c = 0; c < 8; ++c) {
    const int64_t node = ((int64_t)(c0[0] + (c & 1)) * ny + (c0[1] + ((c >> 1) & 1))) * nz +
                         (c0[2] + ((c >> 2) & 1));
    if (a.fp64) {
      const double *p = (const double *)a.grid + node * 8;
      for (int q = 0; q < 8; ++q) X[c][q] = p[q];
    } else {
      const float4 *p = (const float4 *)((const float *)a.grid + node * 8);
      float4 v0 = p[0], v1 = p[1];
      X[c][0] = v0.x; X[c][1] = v0.y; X[c][2] = v0.z; X[c][3] = v0.w;
      X[c][4] = v1.x; X[c][5] = v1.y; X[c][6] = v1.z; X[c][7] = v1.w;
    }
  }
  if (!a.coulomb)
    for (int c = 0; c < 8; ++c)
      if (X[c][0] > 1e10) return 1e8;
  double hx[2][2], hy[2][2], hz[2][2];
  hermite(r[0], hx);
  hermite(r[1], hy);
  hermite(r[2], hz);
  // channel -> derivative orders (dx, dy, dz)
  const int dd[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                        {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  double v = 0.0;
  for (int c = 0; c < 8; ++c) {
    const int cx0 = c & 1, cy0 = (c >> 1) & 1, cz0 = (c >> 2) & 1;
    for (int q = 0; q < 8; ++q)
      v += X[c][q] * (hx[cx0][dd[q][0]] * hy[cy0][dd[q][1]] * hz[cz0][dd[q][2]]);
  }
  return a.coulomb ? a.charge * v : v;
}

__global__ __launch_bounds__(256) void k_interp(const InterpArgs a, const double *__restrict__ xyz,
                                                int64_t n, double *__restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = interp_at(a, xyz[3 * t], xyz[3 * t + 1], xyz[3 * t + 2]);
}

// Site energies from tabulated grids (FrameworkModel GRID): the VDW grid of
// the site's pseudo-atom and/or the Coulomb grid times its charge
// (src/framework_energy.c:73-76, 197-200).  merge keeps the overlap flags a
// pair kernel already wrote for the other term of the same sites.
__global__ __launch_bounds__(256) void k_site_grid(const InterpArgs gv, const InterpArgs gc, int has_v,
                                                   int has_c, const SiteList s, const EnergyOut o,
                                                   int merge) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= site_count(s)) return;
  const int64_t k = site_phys(s, i);
  const double px = s.x[k], py = s.y[k], pz = s.z[k];
  bool flag = false;
  if (has_v) {
    const double uv = interp_at(gv, px, py, pz);
    if (o.u_vdw) o.u_vdw[k] = uv;
    flag = flag || uv >= kOverlap;
  }
  if (has_c) {
    const double uc = interp_at(gc, px, py, pz);
    if (o.u_cc) o.u_cc[k] = uc;
    flag = flag || uc >= kOverlap;
  }
  if (o.overlap) o.overlap[k] = (uint8_t)((merge && o.overlap[k]) || flag);
}

// Guest-guest energies at finite loading (CalculateInterVDWEnergyAdsorbate-
// AtPosition, CalculateInterChargeEnergyAdsorbateAtPosition, src/inter_energy.c:
// 321-520): every atom of every adsorbate molecule but the excluded one, in
// molecule and atom order, FP64 in the reference's operation order.  The
// atoms arrive wave-uniform, one scalar load each.
template <int BCK>
__global__ __launch_bounds__(256) void k_site_ads(const AdsArgs a, const SiteList s, double *__restrict__ out_v,
                                                  double *__restrict__ out_c) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = site_count(s);
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;
  const bool live = i < n;
  const int64_t k = site_phys(s, live ? i : 0);
  const double px = s.x[k], py = s.y[k], pz = s.z[k];
  int ex = a.exclude;
  if (a.excl) {
    const int64_t col = s.in_stride > 0 ? k % s.in_stride : k;
    ex = a.excl[a.col_map ? a.col_map[col] : col];
  }
  double uv = 0.0, uc = 0.0;
  for (int j = 0; j < a.n; ++j) {
    if (a.mol[j] == ex) continue;
    const AtomRec64 r = a.rec[j];
    double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
    mi_exact<BCK>(a.cell, dx, dy, dz);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (a.do_vdw && rr < a.rc2v) uv += lj_exact(r.c4, r.s2, r.shift, rr);
    if (a.do_cc && rr < a.rc2c) {
      const double rd = sqrt(rr);
      uc += a.cqa * r.q * (erfc(a.alpha * rd) / rd);
    }
  }
  if (live) {
    if (out_v) out_v[k] = uv;
    if (out_c) out_c[k] = uc;
  }
}

void launch_site_ads(const AdsArgs &a, const SiteList &s, double *u_vdw, double *u_cc, hipStream_t st) {
  const unsigned nb = nblk(s.n, 256);
  if (!nb) return;
  if (a.cell.bc == 1)
    hipLaunchKernelGGL(k_site_ads<1>, dim3(nb), dim3(256), 0, st, a, s, u_vdw, u_cc);
  else if (a.cell.upper)
    hipLaunchKernelGGL(k_site_ads<2>, dim3(nb), dim3(256), 0, st, a, s, u_vdw, u_cc);
  else
    hipLaunchKernelGGL(k_site_ads<3>, dim3(nb), dim3(256), 0, st, a, s, u_vdw, u_cc);
}

void launch_site_grid(const InterpArgs *gv, const InterpArgs *gc, const SiteList &s, const EnergyOut &o,
                      int merge, hipStream_t st) {
  const unsigned nb = nblk(s.n, 256);
  if (!nb) return;
  InterpArgs z;
  memset(&z, 0, sizeof(z));
  hipLaunchKernelGGL(k_site_grid, dim3(nb), dim3(256), 0, st, gv ? *gv : z, gc ? *gc : z, gv ? 1 : 0,
                     gc ? 1 : 0, s, o, merge);
}

void launch_interp(const InterpArgs &a, const double *xyz, int64_t n, double *out, hipStream_t st) {
  unsigned nb = nblk(n, 256);
  if (nb) hipLaunchKernelGGL(k_interp, dim3(nb), dim3(256), 0, st, a, xyz, n, out);
}

// ====================================================================
// Geometric surface area (src/mc_moves.c:16434-16520): one lane per
// sampled point, the point on its atom's sphere from the counter stream of
// (seed, point) with RandomNumberOnUnitSphere's rejection (src/utils.c:
// 688-708); the overlap test of CheckSurfaceAreaOverlap (src/framework.c:
// 11649-11690) in FP64 with the reference's minimum image, the atoms
// wave-uniform, a wave leaving the loop once all its points overlap.
// ====================================================================
template <int BCK>
__global__ __launch_bounds__(256) void k_surface_area(const SurfaceArgs a,
                                                      unsigned long long *__restrict__ counts) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)a.n * a.samples;
  if ((int64_t)blockIdx.x * blockDim.x >= total) return;
  const bool live = t < total;
  const int64_t tt = live ? t : 0;
  const int i = (int)(tt / a.samples);
  const uint64_t key = stream_key(a.seed, (uint64_t)tt);
  uint64_t c = 0;
  double r1, r2, rsq;
  int guard = 0;
  do {
    r1 = 2.0 * stream_uniform(key, c++) - 1.0;
    r2 = 2.0 * stream_uniform(key, c++) - 1.0;
    rsq = r1 * r1 + r2 * r2;
  } while (rsq >= 1.0 && ++guard < 4096);
  const double rh = 2.0 * sqrt(1.0 - rsq);
  const double vx = r1 * rh, vy = r2 * rh, vz = 1.0 - 2.0 * rsq;
  const double4 ai = a.atoms[i];
  const double px = ai.x + vx * ai.w, py = ai.y + vy * ai.w, pz = ai.z + vz * ai.w;
  bool hit = !live || blocked_pocket(a.pockets, px, py, pz);
  for (int j = 0; j < a.n; ++j) {
    if (__all(hit)) break;
    if (hit || j == i) continue;
    const double4 b = a.atoms[j];
    double dx = px - b.x, dy = py - b.y, dz = pz - b.z;
    mi_exact<BCK>(a.cell, dx, dy, dz);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < b.w * b.w) hit = true;
  }
  if (live && !hit) atomicAdd(counts + i, 1ULL);
}

void launch_surface_area(const SurfaceArgs &a, unsigned long long *counts, hipStream_t st) {
  const int64_t total = (int64_t)a.n * a.samples;
  const unsigned nb = nblk(total, 256);
  if (!nb) return;
  if (a.cell.bc == 1)
    hipLaunchKernelGGL(k_surface_area<1>, dim3(nb), dim3(256), 0, st, a, counts);
  else if (a.cell.upper)
    hipLaunchKernelGGL(k_surface_area<2>, dim3(nb), dim3(256), 0, st, a, counts);
  else
    hipLaunchKernelGGL(k_surface_area<3>, dim3(nb), dim3(256), 0, st, a, counts);
}

// ====================================================================
// plain streaming copy: calibrates hardware byte counters against a
// kernel whose traffic is known (bytes read = bytes written = size)
// ====================================================================
__global__ __launch_bounds__(256) void k_stream_copy(const float4 *__restrict__ src,
                                                     float4 *__restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_stream_copy(const void *src, void *dst, int64_t bytes, hipStream_t st) {
  const int64_t n = bytes / 16;
  hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(256), 0, st, (const float4 *)src, (float4 *)dst, n);
}

// ====================================================================
// Flexible chains: CBMC growth with the internal MC scheme
// (src/cbmc.c:1480-2482, 2762-2849; src/internal_energy.c:48-70,
// 306-329, 372-490, 1168-1340, 2177-2330; src/utils.c:640-990).
// FP64 throughout, no contraction, the reference's operation order; the
// draws come from per-purpose sub-streams of the insertion's counter
// stream.
// ====================================================================
namespace chain {

struct Draw {
  uint64_t key, ctr;
  __device__ double next() { return stream_uniform(key, ctr++); }
};

// RandomGaussianNumber (src/utils.c:640-651)
__device__ double gauss(Draw &d) {
#pragma clang fp contract(off)
  double ran1, ran2, r2;
  do {
    ran1 = 2.0 * d.next() - 1.0;
    ran2 = 2.0 * d.next() - 1.0;
    r2 = ran1 * ran1 + ran2 * ran2;
  } while ((r2 > 1.0) || (r2 == 0.0));
  return ran2 * sqrt(-2.0 * log(r2) / r2);
}

// RandomNumberOnUnitSphere (src/utils.c:688-708)
__device__ void sphere(Draw &d, double v[3]) {
#pragma clang fp contract(off)
  double ran1, ran2, ransq;
  do {
    ran1 = 2.0 * d.next() - 1.0;
    ran2 = 2.0 * d.next() - 1.0;
    ransq = ran1 * ran1 + ran2 * ran2;
  } while (ransq >= 1.0);
  const double ranh = 2.0 * sqrt(1.0 - ransq);
  v[0] = ran1 * ranh;
  v[1] = ran2 * ranh;
  v[2] = 1.0 - 2.0 * ransq;
}

// RandomNumberOnCone (src/utils.c:709-735)
__device__ void cone(Draw &d, const double vin[3], double theta, double v[3]) {
#pragma clang fp contract(off)
  const double cosin = cos(theta);
  double ran1 = 1.0 / sqrt(vin[0] * vin[0] + vin[1] * vin[1] + vin[2] * vin[2]);
  const double xprev = vin[0] * ran1, yprev = vin[1] * ran1, zprev = vin[2] * ran1;
  sphere(d, v);
  ran1 = v[0] * xprev + v[1] * yprev + v[2] * zprev;
  v[0] = v[0] - ran1 * xprev;
  v[1] = v[1] - ran1 * yprev;
  v[2] = v[2] - ran1 * zprev;
  ran1 = sqrt((1.0 - cosin * cosin) / (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
  v[0] = v[0] * ran1 + cosin * xprev;
  v[1] = v[1] * ran1 + cosin * yprev;
  v[2] = v[2] * ran1 + cosin * zprev;
  ran1 = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] = v[0] * ran1;
  v[1] = v[1] * ran1;
  v[2] = v[2] * ran1;
}

// Perpendicular (src/utils.c:789-832)
__device__ void perpendicular(const double a[3], const double b[3], double v[3]) {
#pragma clang fp contract(off)
  double fx = fabs(b[1] * a[2] - a[1] * b[2]);
  const double fy = fabs(b[0] * a[2] - a[0] * b[2]);
  const double fz = fabs(b[1] * a[0] - a[1] * b[0]);
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
  const double f = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] *= f;
  v[1] *= f;
  v[2] *= f;
}

// RotationAroundXYZAxis (src/utils.c:945-975): the matrix of a rotation by
// theta about the unit axis v, applied as the reference applies it
struct Rot {
  double r[3][3];
  __device__ Rot(const double v[3], double theta) {
#pragma clang fp contract(off)
    const double c = cos(theta);
    const double w = 1.0 - c;
    double s = sqrt(1.0 - c * c);
    if (theta < 0.0) s = -s;
    r[0][0] = (v[0]) * (v[0]) * w + c;
    r[0][1] = (v[0]) * (v[1]) * w + (v[2]) * s;
    r[0][2] = (v[0]) * (v[2]) * w - (v[1]) * s;
    r[1][0] = (v[0]) * (v[1]) * w - (v[2]) * s;
    r[1][1] = (v[1]) * (v[1]) * w + c;
    r[1][2] = (v[1]) * (v[2]) * w + (v[0]) * s;
    r[2][0] = (v[0]) * (v[2]) * w + (v[1]) * s;
    r[2][1] = (v[1]) * (v[2]) * w - (v[0]) * s;
    r[2][2] = (v[2]) * (v[2]) * w + c;
  }
  __device__ void apply(double p[3]) const {
#pragma clang fp contract(off)
    const double x = p[0] * r[0][0] + p[1] * r[0][1] + p[2] * r[0][2];
    const double y = p[0] * r[1][0] + p[1] * r[1][1] + p[2] * r[1][2];
    const double z = p[0] * r[2][0] + p[1] * r[2][1] + p[2] * r[2][2];
    p[0] = x;
    p[1] = y;
    p[2] = z;
  }
};

// CalculateAngle (src/internal_energy.c:306-329)
__device__ double angle(const double a[3], const double b[3], const double c[3]) {
#pragma clang fp contract(off)
  double ab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  const double rab = 1.0 / sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
  ab[0] *= rab;
  ab[1] *= rab;
  ab[2] *= rab;
  double bc[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  const double rbc = 1.0 / sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
  bc[0] *= rbc;
  bc[1] *= rbc;
  bc[2] *= rbc;
  return acos(ab[0] * bc[0] + ab[1] * bc[1] + ab[2] * bc[2]);
}

// CalculateBondEnergy, HARMONIC_BOND (src/internal_energy.c:372-395); FIXED_BOND is 0
__device__ double bond_u(int kind, double p0, double p1, const double a[3], const double b[3]) {
#pragma clang fp contract(off)
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (kind == 1) return 0.5 * p0 * ((r - p1) * (r - p1));
  return 0.0;
}

// CalculateBendEnergy, HARMONIC_BEND (src/internal_energy.c:1225-1268)
__device__ double bend_u(double p0, double p1, const double a[3], const double b[3], const double c[3]) {
#pragma clang fp contract(off)
  double ab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  const double rab = sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
  ab[0] /= rab;
  ab[1] /= rab;
  ab[2] /= rab;
  double bc[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  const double rbc = sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
  bc[0] /= rbc;
  bc[1] /= rbc;
  bc[2] /= rbc;
  double ct = ab[0] * bc[0] + ab[1] * bc[1] + ab[2] * bc[2];
  const double lo = (-1.0 > ct) ? -1.0 : ct;
  ct = (1.0 <= lo) ? 1.0 : lo;
  const double th = acos(ct);
  return 0.5 * p0 * ((th - p1) * (th - p1));
}

// CalculateTorsionEnergy, TRAPPE_DIHEDRAL (src/internal_energy.c:2177-2230, 2324-2332)
__device__ double torsion_u(const double *p, const double a[3], const double b[3], const double c[3],
                            const double d[3]) {
#pragma clang fp contract(off)
  const double Dab[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
  double Dbc[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
  const double rbc = sqrt(Dbc[0] * Dbc[0] + Dbc[1] * Dbc[1] + Dbc[2] * Dbc[2]);
  Dbc[0] /= rbc;
  Dbc[1] /= rbc;
  Dbc[2] /= rbc;
  const double Dcd[3] = {d[0] - c[0], d[1] - c[1], d[2] - c[2]};
  const double dot_ab = Dab[0] * Dbc[0] + Dab[1] * Dbc[1] + Dab[2] * Dbc[2];
  const double dot_cd = Dcd[0] * Dbc[0] + Dcd[1] * Dbc[1] + Dcd[2] * Dbc[2];
  double dr[3] = {Dab[0] - dot_ab * Dbc[0], Dab[1] - dot_ab * Dbc[1], Dab[2] - dot_ab * Dbc[2]};
  const double r = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
  dr[0] /= r;
  dr[1] /= r;
  dr[2] /= r;
  double ds[3] = {Dcd[0] - dot_cd * Dbc[0], Dcd[1] - dot_cd * Dbc[1], Dcd[2] - dot_cd * Dbc[2]};
  const double s = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  ds[0] /= s;
  ds[1] /= s;
  ds[2] /= s;
  double cp = dr[0] * ds[0] + dr[1] * ds[1] + dr[2] * ds[2];
  const double m = fabs(cp) <= 1.0 ? fabs(cp) : 1.0;
  cp = cp >= 0.0 ? m : -m;
  return p[0] + (1.0 + cp) * (p[1] + p[3] - 2.0 * (cp - 1.0) * (p[2] - 2.0 * p[3] * cp));
}

// GenerateBondlength (src/internal_energy.c:48-70, 185-187)
__device__ double bond_length(Draw &d, double beta, int kind, double p0, double p1) {
#pragma clang fp contract(off)
  if (kind != 1) return p0;
  const double ran1 = 1.0 / sqrt(beta * p0);
  const double ran2 = 1.0 / ((p1 + 3.0 * ran1) * (p1 + 3.0 * ran1));
  double bl;
  do
    bl = p1 + ran1 * gauss(d);
  while ((d.next() > (bl * bl * ran2)) || (bl <= 0.0));
  return bl;
}

// The positions one lane works on: the beads placed by the step (0..3), the
// current and previous beads in registers, and any other placed atom read
// from the grown molecule.  Codes come from the step table (wave-uniform),
// so get() resolves them with scalar branches and constant register indices.
template <int NP>
struct Local {
  double p[NP][3];
  double c[3], v[3];
  const ChainWork *w;
  int64_t slot;
  __device__ __forceinline__ void get(int code, double out[3]) const {
    if (code >= kLocGlobal) {
      const int64_t o = (int64_t)(code - kLocGlobal) * w->cap + slot;
      out[0] = w->px[o];
      out[1] = w->py[o];
      out[2] = w->pz[o];
    } else if (code == kLocCur) {
      out[0] = c[0];
      out[1] = c[1];
      out[2] = c[2];
    } else if (code == kLocPrev) {
      out[0] = v[0];
      out[1] = v[1];
      out[2] = v[2];
    } else {
#pragma unroll
      for (int j = 0; j < NP; ++j)
        if (code == j) {
          out[0] = p[j][0];
          out[1] = p[j][1];
          out[2] = p[j][2];
        }
    }
  }
  __device__ __forceinline__ void load(int atom, double out[3]) const {
    const int64_t o = (int64_t)atom * w->cap + slot;
    out[0] = w->px[o];
    out[1] = w->py[o];
    out[2] = w->pz[o];
  }
};

// register arrays indexed by a per-lane value: compare-and-select over the
// constant indices keeps them out of scratch memory
template <int N>
__device__ __forceinline__ void pick3(const double (&a)[N][3], int j, double out[3]) {
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (j == q) {
      out[0] = a[q][0];
      out[1] = a[q][1];
      out[2] = a[q][2];
    }
}
template <int N>
__device__ __forceinline__ void put3(double (&a)[N][3], int j, const double in[3]) {
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (j == q) {
      a[q][0] = in[0];
      a[q][1] = in[1];
      a[q][2] = in[2];
    }
}
template <int N>
__device__ __forceinline__ double pick1(const double (&a)[N], int j) {
  double r = 0.0;
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (j == q) r = a[q];
  return r;
}
template <int N>
__device__ __forceinline__ void put1(double (&a)[N], int j, double v) {
#pragma unroll
  for (int q = 0; q < N; ++q)
    if (j == q) a[q] = v;
}

template <class LT>
__device__ __forceinline__ double bond_j(const ChainStep &S, const LT &L, int j) {
  double a[3], b[3];
  L.get(S.bond_a[j], a);
  L.get(S.bond_b[j], b);
  return bond_u(S.bond_kind[j], S.bond_p0[j], S.bond_p1[j], a, b);
}
template <class LT>
__device__ __forceinline__ double bend_j(const ChainStep &S, const LT &L, int j) {
  double a[3], b[3], c[3];
  L.get(S.bend_a[j], a);
  L.get(S.bend_b[j], b);
  L.get(S.bend_c[j], c);
  return bend_u(S.bend_p0[j], S.bend_p1[j], a, b, c);
}
template <class LT>
__device__ double sum_bonds(const ChainStep &S, const LT &L) {
#pragma clang fp contract(off)
  double u = 0.0;
  for (int j = 0; j < S.n_bonds; ++j) u += bond_j(S, L, j);
  return u;
}
template <class LT>
__device__ double sum_bends(const ChainStep &S, const LT &L) {
#pragma clang fp contract(off)
  double u = 0.0;
  for (int j = 0; j < S.n_bends; ++j) u += bend_j(S, L, j);
  return u;
}
template <class LT>
__device__ double torsion_j(const ChainStep &S, const LT &L, int j) {
  double a[3], b[3], c[3], d[3];
  L.get(S.tors_a[j], a);
  L.get(S.tors_b[j], b);
  L.get(S.tors_c[j], c);
  L.get(S.tors_d[j], d);
  return torsion_u(S.tors_p[j], a, b, c, d);
}
template <class LT>
__device__ double sum_torsions(const ChainStep &S, const LT &L) {
#pragma clang fp contract(off)
  double u = 0.0;
  for (int j = 0; j < S.n_tors; ++j) u += torsion_j(S, L, j);
  return u;
}
// intra-molecular LJ of the step's pairs (src/cbmc.c:1294-1311): plain
// distance, cut-off CutOffVDW, scaling times PotentialValue
template <class LT>
__device__ double sum_intra(const ChainStep &S, const LT &L, double rc2) {
#pragma clang fp contract(off)
  double u = 0.0;
  for (int j = 0; j < S.n_intra; ++j) {
    double a[3], b[3];
    L.get(S.intra_a[j], a);
    L.get(S.intra_b[j], b);
    const double rr = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
    if (rr < rc2) u += S.intra_scal[j] * lj_exact(S.intra_c4[j], S.intra_s2[j], S.intra_shift[j], rr);
  }
  return u;
}

// A torsion whose only moving atom (one placed by the step) is an end atom:
// the terms of CalculateTorsionEnergy that involve the three fixed atoms are
// computed once per trial (the same operations, so the same values), and
// each rotation evaluates the moving half only.
struct TorsionHalf {
  int mode;      // 0: generic, 1: a moves, 2: d moves
  int mover;     // local slot of the moving atom
  double dbc[3]; // normalised c - b
  double fix[3]; // normalised perpendicular part of the fixed end
  double q[3];   // the fixed atom adjacent to the moving one (b for a, c for d)
};

template <class LT>
__device__ __forceinline__ void torsion_prepare(const ChainStep &S, const LT &L, int j, TorsionHalf &h) {
#pragma clang fp contract(off)
  const int ca = S.tors_a[j], cb = S.tors_b[j], cc = S.tors_c[j], cd = S.tors_d[j];
  const bool ma = ca < kLocCur, mb = cb < kLocCur, mc = cc < kLocCur, md = cd < kLocCur;
  h.mode = 0;
  if (mb || mc || (ma && md) || (!ma && !md)) return;
  double b[3], c[3];
  L.get(cb, b);
  L.get(cc, c);
  h.dbc[0] = c[0] - b[0];
  h.dbc[1] = c[1] - b[1];
  h.dbc[2] = c[2] - b[2];
  const double rbc = sqrt(h.dbc[0] * h.dbc[0] + h.dbc[1] * h.dbc[1] + h.dbc[2] * h.dbc[2]);
  h.dbc[0] /= rbc;
  h.dbc[1] /= rbc;
  h.dbc[2] /= rbc;
  double e[3], f[3];
  if (md) {
    // fixed a: Dab and dr
    h.mode = 2;
    h.mover = cd;
    L.get(ca, e);
    f[0] = e[0] - b[0];
    f[1] = e[1] - b[1];
    f[2] = e[2] - b[2];
    h.q[0] = c[0];
    h.q[1] = c[1];
    h.q[2] = c[2];
  } else {
    // fixed d: Dcd and ds
    h.mode = 1;
    h.mover = ca;
    L.get(cd, e);
    f[0] = e[0] - c[0];
    f[1] = e[1] - c[1];
    f[2] = e[2] - c[2];
    h.q[0] = b[0];
    h.q[1] = b[1];
    h.q[2] = b[2];
  }
  const double dot = f[0] * h.dbc[0] + f[1] * h.dbc[1] + f[2] * h.dbc[2];
  h.fix[0] = f[0] - dot * h.dbc[0];
  h.fix[1] = f[1] - dot * h.dbc[1];
  h.fix[2] = f[2] - dot * h.dbc[2];
  const double r = sqrt(h.fix[0] * h.fix[0] + h.fix[1] * h.fix[1] + h.fix[2] * h.fix[2]);
  h.fix[0] /= r;
  h.fix[1] /= r;
  h.fix[2] /= r;
}

// CalculateTorsionEnergy from a prepared half (TRAPPE_DIHEDRAL)
__device__ __forceinline__ double torsion_half(const double *par, const TorsionHalf &h, const double m[3]) {
#pragma clang fp contract(off)
  // the moving end: Dab = a - b (mode 1) or Dcd = d - c (mode 2)
  const double g[3] = {m[0] - h.q[0], m[1] - h.q[1], m[2] - h.q[2]};
  const double dot = g[0] * h.dbc[0] + g[1] * h.dbc[1] + g[2] * h.dbc[2];
  double u[3] = {g[0] - dot * h.dbc[0], g[1] - dot * h.dbc[1], g[2] - dot * h.dbc[2]};
  const double r = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  u[0] /= r;
  u[1] /= r;
  u[2] /= r;
  // CosPhi = dr . ds with dr from the a end
  double cp = (h.mode == 1) ? (u[0] * h.fix[0] + u[1] * h.fix[1] + u[2] * h.fix[2])
                            : (h.fix[0] * u[0] + h.fix[1] * u[1] + h.fix[2] * u[2]);
  const double mm = fabs(cp) <= 1.0 ? fabs(cp) : 1.0;
  cp = cp >= 0.0 ? mm : -mm;
  return par[0] + (1.0 + cp) * (par[1] + par[3] - 2.0 * (cp - 1.0) * (par[2] - 2.0 * par[3] * cp));
}

__device__ __forceinline__ uint64_t insertion_key(const ChainArgs &a, const ChainWork &w, int64_t slot) {
  return stream_key(a.seed, (uint64_t)(a.first + w.active[slot]));
}

__device__ __forceinline__ double wave_max(double v) {
  for (int s = 32; s > 0; s >>= 1) v = fmax(v, __shfl_xor(v, s));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s);
  return v;
}
// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ double wave_scan(double v, int lane) {
  for (int s = 1; s < 64; s <<= 1) {
    const double u = __shfl_up(v, s);
    if (lane >= s) v += u;
  }
  return v;
}

} // namespace chain

// HandleFirstBead selection (src/cbmc.c:695-795) for chains: the accepted
// insertions get a slot of the growing set.
__global__ __launch_bounds__(256) void k_chain_fb_select(const ChainArgs a, const ChainWork w) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const int k = a.k_fb;
  const double beta = a.beta;
  bool any = false;
  double W1 = 0.0;
  for (int t = 0; t < k; ++t) {
    const double uv = w.fb_uv[t * w.cap + i];
    if (uv < kOverlap) {
      any = true;
      W1 += exp(-beta * (uv + 0.0 + 0.0));
    }
  }
  w.W[i] = 0.0;
  w.dU[i] = 0.0;
  double *rec = w.rec ? w.rec + i * kCrecN : nullptr;
  if (rec)
    for (int f = 0; f < kCrecN; ++f) rec[f] = 0.0;
  if (!any) {
    if (rec) rec[kCrecOverlap] = 1.0;
    return;
  }
  const uint64_t key = stream_key(a.seed, (uint64_t)(a.first + i));
  const double usel = stream_uniform(key, (uint64_t)(3 * k));
  auto bf = [&](int t, double &b) -> bool {
    const double uv = w.fb_uv[t * w.cap + i];
    if (uv >= kOverlap) return false;
    b = -beta * (uv + 0.0 + 0.0);
    return true;
  };
  const int sel = select_trial(bf, k, usel);
  W1 /= (double)k;
  if (rec) rec[kCrecW1] = W1;
  if (W1 < kMinRosenbluth) {
    if (rec) {
      rec[kCrecOverlap] = 1.0;
      rec[kCrecStage] = 1.0;
    }
    return;
  }
  const int64_t slot = wave_slot(w.n_active);
  w.active[slot] = (int32_t)i;
  w.alive[slot] = 1;
  w.sW[slot] = W1;
  w.W1[slot] = W1;
  w.uhost[slot] = w.fb_uv[sel * w.cap + i];
  w.uintra[slot] = 0.0;
  w.ubond[slot] = 0.0;
  w.ubend[slot] = 0.0;
  w.utors[slot] = 0.0;
  const int64_t o = (int64_t)a.start * w.cap + slot;
  w.px[o] = w.fbx[sel * w.cap + i];
  w.py[o] = w.fby[sel * w.cap + i];
  w.pz[o] = w.fbz[sel * w.cap + i];
}

// Bond lengths of a step (GenerateBondlength); for a step with a previous
// bead the internal Metropolis walk of GenerateTrialOrientationsMCScheme
// (src/cbmc.c:1580-2350) without a stored configuration, on trial 0.
template <int NPM>
__global__ __launch_bounds__(64) void k_chain_grow(const ChainArgs a, const ChainWork w) {
#pragma clang fp contract(off)
  constexpr int NBE = NPM == 1 ? 2 : kChainBends;
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= (int64_t)*w.n_active || !w.alive[slot]) return;
  const ChainStep &S = *a.st;
  const double beta = a.beta;
  chain::Draw d{chain_substream(chain::insertion_key(a, w, slot), chain_tag(a.step, 0)), 0};
  chain::Local<NPM> L;
  L.w = &w;
  L.slot = slot;
  L.load(S.cur, L.c);
  if (S.n_prev == 0) {
    double bl0 = 0.0;
    for (int j = 0; j < S.n_bonds; ++j) {
      const double b = chain::bond_length(d, beta, S.bond_kind[j], S.bond_p0[j], S.bond_p1[j]);
      if (j == 0) bl0 = b;
    }
    w.bl[slot] = bl0;
    return;
  }
  L.load(S.prev, L.v);
  for (int j = 0; j < S.n_bonds; ++j) (void)chain::bond_length(d, beta, S.bond_kind[j], S.bond_p0[j], S.bond_p1[j]);
  const double *cur = L.c;
  double pv[3] = {L.v[0] - cur[0], L.v[1] - cur[1], L.v[2] - cur[2]};
  const double nrm = sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
  pv[0] /= nrm;
  pv[1] /= nrm;
  pv[2] /= nrm;
  const int nb = S.n_bonds, nbe = S.n_bends, np = S.n_place;
#pragma unroll
  for (int jj = 0; jj < NPM; ++jj)
    if (jj < np) {
      double v[3];
      chain::cone(d, pv, 120 * (M_PI / 180.0), v);
      for (int q = 0; q < 3; ++q) L.p[jj][q] = cur[q] + S.ref_len[jj] * v[q];
    }
  double eobond[NPM], enbond[NPM], eobend[NBE], enbend[NBE];
#pragma unroll
  for (int j = 0; j < NPM; ++j)
    if (j < nb) eobond[j] = enbond[j] = chain::bond_j(S, L, j);
#pragma unroll
  for (int j = 0; j < NBE; ++j)
    if (j < nbe) eobend[j] = enbend[j] = chain::bend_j(S, L, j);
  const int trialmove = 4 * a.moves * np;
  // no chiral centres: two branches swap places with probability 1/2
  // (src/cbmc.c:1814-1838); steps placing one bead have a single bond
  if constexpr (NPM >= 2) {
    if ((nb >= 2) && (d.next() < 0.5)) {
      double cA[3], cC[3];
      for (int q = 0; q < 3; ++q) {
        cA[q] = L.p[0][q] - cur[q];
        cC[q] = L.p[1][q] - cur[q];
      }
      double tmp = sqrt(cA[0] * cA[0] + cA[1] * cA[1] + cA[2] * cA[2]) / sqrt(cC[0] * cC[0] + cC[1] * cC[1] + cC[2] * cC[2]);
      for (int q = 0; q < 3; ++q) L.p[0][q] = cur[q] + tmp * cC[q];
      tmp = 1.0 / tmp;
      for (int q = 0; q < 3; ++q) L.p[1][q] = cur[q] + tmp * cA[q];
  #pragma unroll
      for (int j = 0; j < NPM; ++j)
        if (j < nb) eobond[j] = enbond[j] = chain::bond_j(S, L, j);
  #pragma unroll
      for (int j = 0; j < NBE; ++j)
        if (j < nbe) eobend[j] = enbend[j] = chain::bend_j(S, L, j);
    }
  }
  for (int jjj = 0; jjj < trialmove; ++jjj) {
    int choise;
    if (nbe == 1)
      choise = d.next() < 0.3 ? 1 : 2;
    else if (jjj < 20)
      choise = 3;
    else {
      const double tmp = d.next();
      choise = tmp < 0.2 ? 1 : (tmp < 0.6 ? 2 : 3);
    }
#pragma unroll
    for (int j = 0; j < NPM; ++j)
      if (j < nb) enbond[j] = eobond[j];
#pragma unroll
    for (int j = 0; j < NBE; ++j)
      if (j < nbe) enbend[j] = eobend[j];
    double newbond = 0.0, newbend = 0.0;
    if (choise == 1) {
      if (nb > 0) {
        int jj = (int)(d.next() * (double)nb);
        if (jj >= nb) jj = nb - 1; // u = 1 (closed interval) would index past the table
        if (S.bond_kind[jj] != 0) {
          const int jo = S.bond_branch[jj];
          double store[3];
          chain::pick3(L.p, jo, store);
          const double vec[3] = {store[0] - cur[0], store[1] - cur[1], store[2] - cur[2]};
          const double blo = sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
          const double bln = blo + (2.0 * d.next() - 1.0) * S.max_bond[jo];
          double np3[3];
          for (int q = 0; q < 3; ++q) np3[q] = cur[q] + (bln / blo) * vec[q];
          chain::put3(L.p, jo, np3);
          const double e = chain::bond_j(S, L, jj);
          chain::put1(enbond, jj, e);
          newbond += e - chain::pick1(eobond, jj);
          if (d.next() < ((bln / blo) * (bln / blo)) * exp(-beta * (newbond + 0.0 + 0.0 + 0.0)))
            chain::put1(eobond, jj, e);
          else
            chain::put3(L.p, jo, store);
        }
      }
    } else if (choise == 2) {
      if (np > 0) {
        int jo = (int)(d.next() * (double)np);
        if (jo >= np) jo = np - 1;
        double store[3];
        chain::pick3(L.p, jo, store);
        double cord[3] = {store[0] - cur[0], store[1] - cur[1], store[2] - cur[2]};
        const double sin_old = sin(chain::angle(L.v, cur, store));
        const double dr[3] = {cur[0] - store[0], cur[1] - store[1], cur[2] - store[2]};
        double perpen[3];
        chain::perpendicular(dr, pv, perpen);
        const chain::Rot R(perpen, (2.0 * d.next() - 1.0) * S.max_bend);
        R.apply(cord);
        double np3[3];
        for (int q = 0; q < 3; ++q) np3[q] = cur[q] + cord[q];
        chain::put3(L.p, jo, np3);
        const double sin_new = sin(chain::angle(L.v, cur, np3));
#pragma unroll
        for (int j = 0; j < NBE; ++j)
          if (j < nbe) {
            enbend[j] = chain::bend_j(S, L, j);
            newbend += enbend[j] - eobend[j];
          }
        if (d.next() < ((sin_new / sin_old) * exp(-beta * (newbend + 0.0 + 0.0 + 0.0)))) {
#pragma unroll
          for (int j = 0; j < NBE; ++j)
            if (j < nbe) eobend[j] = enbend[j];
        } else
          chain::put3(L.p, jo, store);
      }
    } else {
      if (np > 0) {
        int jo = (int)(d.next() * (double)np);
        if (jo >= np) jo = np - 1;
        double store[3];
        chain::pick3(L.p, jo, store);
        double cord[3] = {store[0] - cur[0], store[1] - cur[1], store[2] - cur[2]};
        const chain::Rot R(pv, (2.0 * d.next() - 1.0) * S.max_cone);
        R.apply(cord);
        double np3[3];
        for (int q = 0; q < 3; ++q) np3[q] = cur[q] + cord[q];
        chain::put3(L.p, jo, np3);
#pragma unroll
        for (int j = 0; j < NBE; ++j)
          if (j < nbe) {
            enbend[j] = chain::bend_j(S, L, j);
            newbend += enbend[j] - eobend[j];
          }
        if (d.next() < exp(-beta * (newbend + 0.0 + 0.0 + 0.0))) {
#pragma unroll
          for (int j = 0; j < NBE; ++j)
            if (j < nbe) eobend[j] = enbend[j];
        } else
          chain::put3(L.p, jo, store);
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < NPM; ++jj)
    if (jj < np) {
      const int64_t o = (int64_t)jj * w.cap + slot;
      w.bx[o] = L.p[jj][0];
      w.by[o] = L.p[jj][1];
      w.bz[o] = L.p[jj][2];
    }
}

// Trial positions of a step, one wave per (slot, trial).  Sphere steps
// (no previous bead): a random direction at the drawn bond length
// (src/cbmc.c:1530-1577).  Otherwise the torsion selection of
// GenerateTrialOrientationsMCScheme (src/cbmc.c:2350-2435): the lanes take
// the k_torsion rotations of the walk's result about the previous-current
// axis (KPER per lane); a wave maximum, scan and ballot select one with
// probability proportional to exp(-beta U_torsion), and RosenbluthTorsion
// averages the first NumberOfTrialPositions of them (the reference's
// normalisation reads the global trial count, src/cbmc.c:542-555).  Then
// the trial's bonded and intra-molecular energies (src/cbmc.c:2437-2480,
// 1294-1311).
template <int KPER, int NPM>
__global__ __launch_bounds__(256) void k_chain_trials(const ChainArgs a, const ChainWork w) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int k = a.k;
  const int64_t slot = wid / k;
  const int t = (int)(wid % k);
  if (slot >= (int64_t)*w.n_active) return;
  const ChainStep &S = *a.st;
  const int64_t cap = w.cap;
  chain::Local<NPM> L;
  L.w = &w;
  L.slot = slot;
  L.load(S.cur, L.c);
  const double *cur = L.c;
  const int np = S.n_place;
  auto store_trial = [&](double rt, double ub, double ube, double ut, double ui) {
#pragma unroll
    for (int jj = 0; jj < NPM; ++jj)
      if (jj < np) {
        const int64_t o = (int64_t)(jj * k + t) * cap + slot;
        w.tx[o] = L.p[jj][0];
        w.ty[o] = L.p[jj][1];
        w.tz[o] = L.p[jj][2];
      }
    const int64_t o = (int64_t)t * cap + slot;
    w.tRT[o] = rt;
    w.tbond[o] = ub;
    w.tbend[o] = ube;
    w.ttors[o] = ut;
    w.tintra[o] = ui;
  };
  if (!w.alive[slot]) {
    // a rejected insertion keeps valid (unused) coordinates in the site lists
    if (lane == 0) {
#pragma unroll
      for (int jj = 0; jj < NPM; ++jj)
        for (int q = 0; q < 3; ++q) L.p[jj][q] = cur[q];
      store_trial(1.0, 0.0, 0.0, 0.0, 0.0);
    }
    return;
  }
  chain::Draw d{chain_substream(chain::insertion_key(a, w, slot), chain_tag(a.step, 2 + t)), 0};
  if (S.n_prev == 0) {
    if (lane == 0) {
      double v[3];
      chain::sphere(d, v);
      const double bl = w.bl[slot];
      for (int q = 0; q < 3; ++q) L.p[0][q] = cur[q] + v[q] * bl;
      store_trial(1.0, chain::bond_j(S, L, 0), 0.0, 0.0, chain::sum_intra(S, L, a.rc2));
    }
    return;
  }
  L.load(S.prev, L.v);
  double pv[3] = {L.v[0] - cur[0], L.v[1] - cur[1], L.v[2] - cur[2]};
  const double nrm = sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
  pv[0] /= nrm;
  pv[1] /= nrm;
  pv[2] /= nrm;
  double base[NPM][3];
#pragma unroll
  for (int jj = 0; jj < NPM; ++jj) {
    const int64_t o = (int64_t)(jj < np ? jj : 0) * cap + slot;
    base[jj][0] = w.bx[o];
    base[jj][1] = w.by[o];
    base[jj][2] = w.bz[o];
    L.p[jj][0] = base[jj][0];
    L.p[jj][1] = base[jj][1];
    L.p[jj][2] = base[jj][2];
  }
  const double beta = a.beta;
  const int kt = a.kt;
  constexpr int kTC = NPM == 1 ? 2 : 4; // torsions whose fixed halves stay in registers
  chain::TorsionHalf th[kTC];
  const bool cached = S.n_tors <= kTC;
#pragma unroll
  for (int j = 0; j < kTC; ++j) {
    th[j].mode = 0;
    if (cached && j < S.n_tors) chain::torsion_prepare(S, L, j, th[j]);
  }
  auto place_at = [&](double ang) {
    const chain::Rot R(pv, ang);
#pragma unroll
    for (int jj = 0; jj < NPM; ++jj)
      if (jj < np) {
        double c[3] = {base[jj][0] - cur[0], base[jj][1] - cur[1], base[jj][2] - cur[2]};
        R.apply(c);
        for (int q = 0; q < 3; ++q) L.p[jj][q] = c[q] + cur[q];
      }
  };
  auto torsions = [&]() -> double {
    if (!cached) return chain::sum_torsions(S, L);
    double u = 0.0;
#pragma unroll
    for (int j = 0; j < kTC; ++j)
      if (j < S.n_tors) {
        if (th[j].mode) {
          double m[3];
          chain::pick3(L.p, th[j].mover, m);
          u += chain::torsion_half(S.tors_p[j], th[j], m);
        } else
          u += chain::torsion_j(S, L, j);
      }
    return u;
  };
  double bf[KPER], e[KPER];
  double largest = -DBL_MAX;
#pragma unroll
  for (int r = 0; r < KPER; ++r) {
    const int kk = r * 64 + lane;
    bf[r] = 0.0;
    if (r * 64 < kt && kk < kt) {
      place_at((2.0 * stream_uniform(d.key, (uint64_t)kk) - 1.0) * M_PI);
      bf[r] = -beta * torsions();
      largest = fmax(largest, bf[r]);
    }
  }
  largest = chain::wave_max(largest);
  // RosenbluthTorsion: the first k of the torsion trials (k <= 64)
  const double rt = chain::wave_sum(lane < k ? exp(bf[0]) : 0.0) / (double)k;
  const double u = stream_uniform(d.key, (uint64_t)kt);
  int sel = 0;
  if (S.n_tors > 0) {
    double total = 0.0;
#pragma unroll
    for (int r = 0; r < KPER; ++r) {
      e[r] = (r * 64 + lane < kt) ? exp(bf[r] - largest) : 0.0;
      if (r * 64 < kt) total += chain::wave_sum(e[r]);
    }
    const double ws = u * total;
    double carry = 0.0;
    sel = -1;
#pragma unroll
    for (int r = 0; r < KPER; ++r) {
      if (r * 64 >= kt || sel >= 0) continue;
      const int kk = r * 64 + lane;
      const double inc = chain::wave_scan(e[r], lane);
      const uint64_t hit = __ballot(kk < kt && carry + inc >= ws);
      if (hit)
        sel = r * 64 + (__ffsll((unsigned long long)hit) - 1);
      else
        carry += __shfl(inc, 63);
    }
    if (sel < 0) sel = kt - 1;
  } else {
    sel = (int)(u * (double)kt);
    if (sel >= kt) sel = kt - 1;
  }
  if (lane != 0) return;
  place_at((2.0 * stream_uniform(d.key, (uint64_t)sel) - 1.0) * M_PI);
  store_trial(rt, chain::sum_bonds(S, L), chain::sum_bends(S, L), torsions(), chain::sum_intra(S, L, a.rc2));
}

// Rosen's selection among the step's trials (src/cbmc.c:2786-2840) with
// ComputeExternalEnergies' overlap rule and LJ biasing (src/cbmc.c:1202-1215,
// 1440-1446).
__global__ __launch_bounds__(256) void k_chain_select(const ChainArgs a, const ChainWork w) {
#pragma clang fp contract(off)
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= (int64_t)*w.n_active || !w.alive[slot]) return;
  const ChainStep &S = *a.st;
  const int k = a.k, np = S.n_place;
  const int64_t cap = w.cap;
  const double beta = a.beta;
  auto host = [&](int t, double &e) -> bool {
    e = 0.0;
    for (int jj = 0; jj < np; ++jj) {
      const double u = w.tuv[(int64_t)(jj * k + t) * cap + slot];
      if (u >= kOverlap) return false;
      e += u;
    }
    return true;
  };
  auto bf = [&](int t, double &b) -> bool {
    double e;
    if (!host(t, e)) return false;
    b = -beta * (e + 0.0 + 0.0 + w.tintra[(int64_t)t * cap + slot]);
    return true;
  };
  bool any = false;
  double weight = 0.0;
  for (int t = 0; t < k; ++t) {
    double b;
    if (bf(t, b)) {
      any = true;
      weight += exp(b);
    }
  }
  const int64_t i = w.active[slot];
  double *rec = w.rec ? w.rec + i * kCrecN : nullptr;
  if (!any) {
    w.alive[slot] = 0;
    if (rec) {
      rec[kCrecOverlap] = 1.0;
      rec[kCrecStage] = 2.0;
      rec[kCrecStep] = (double)a.step;
    }
    return;
  }
  weight /= (double)k;
  const double u = stream_uniform(chain_substream(chain::insertion_key(a, w, slot), chain_tag(a.step, 1)), 0);
  const int iwalk = select_trial(bf, k, u);
  const double W = w.sW[slot] * (weight * w.tRT[(int64_t)iwalk * cap + slot]);
  if (W < kMinRosenbluth) {
    w.alive[slot] = 0;
    if (rec) {
      rec[kCrecOverlap] = 1.0;
      rec[kCrecStage] = 3.0;
      rec[kCrecStep] = (double)a.step;
    }
    return;
  }
  w.sW[slot] = W;
  const int64_t o = (int64_t)iwalk * cap + slot;
  double eh;
  host(iwalk, eh);
  w.ubond[slot] += w.tbond[o];
  w.ubend[slot] += w.tbend[o];
  w.utors[slot] += w.ttors[o];
  w.uintra[slot] += w.tintra[o];
  w.uhost[slot] += eh;
  for (int jj = 0; jj < np; ++jj) {
    const int64_t src = (int64_t)(jj * k + iwalk) * cap + slot;
    const int64_t dst = (int64_t)S.place[jj] * cap + slot;
    w.px[dst] = w.tx[src];
    w.py[dst] = w.ty[src];
    w.pz[dst] = w.tz[src];
  }
}

// WidomAdsorbateMove after the growth (src/mc_moves.c:7880-7933): blocked
// pockets, DeltaU in the reference's order, per-insertion results.
__global__ __launch_bounds__(256) void k_chain_finalize(const ChainArgs a, const ChainWork w) {
#pragma clang fp contract(off)
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= (int64_t)*w.n_active) return;
  const int64_t i = w.active[slot];
  double *rec = w.rec ? w.rec + i * kCrecN : nullptr;
  if (rec) rec[kCrecW1] = w.W1[slot];
  if (!w.alive[slot]) return;
  const int64_t cap = w.cap;
  bool blocked = false;
  for (int j = 0; j < a.n_atoms && !blocked; ++j)
    blocked = blocked_pocket(a.pockets, w.px[j * cap + slot], w.py[j * cap + slot], w.pz[j * cap + slot]);
  if (w.xyz)
    for (int j = 0; j < a.n_atoms; ++j) {
      w.xyz[(i * a.n_atoms + j) * 3 + 0] = w.px[j * cap + slot];
      w.xyz[(i * a.n_atoms + j) * 3 + 1] = w.py[j * cap + slot];
      w.xyz[(i * a.n_atoms + j) * 3 + 2] = w.pz[j * cap + slot];
    }
  if (blocked) {
    if (rec) {
      rec[kCrecStage] = 5.0;
      rec[kCrecUHost] = w.uhost[slot];
    }
    return;
  }
  const double du = w.ubond[slot] + w.ubend[slot] + w.utors[slot] + w.uintra[slot] + w.uhost[slot];
  w.W[i] = w.sW[slot];
  w.dU[i] = du;
  if (rec) {
    rec[kCrecW] = w.sW[slot];
    rec[kCrecDU] = du;
    rec[kCrecUHost] = w.uhost[slot];
    rec[kCrecUIntra] = w.uintra[slot];
    rec[kCrecUBond] = w.ubond[slot];
    rec[kCrecUBend] = w.ubend[slot];
    rec[kCrecUTors] = w.utors[slot];
    rec[kCrecStage] = 4.0;
  }
}

void launch_chain_fb_select(const ChainArgs &a, const ChainWork &w, hipStream_t st) {
  hipLaunchKernelGGL(k_chain_fb_select, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a, w);
}
// steps that place one bead with at most one bond and two bends (every step
// of an unbranched chain) run the one-bead instances
static bool one_bead(const ChainStep &h) { return h.n_place == 1 && h.n_bonds <= 1 && h.n_bends <= 2; }

void launch_chain_grow(const ChainArgs &a, const ChainWork &w, const ChainStep &h, hipStream_t st) {
  const dim3 grid((unsigned)((a.n + 63) / 64));
  if (one_bead(h))
    hipLaunchKernelGGL(k_chain_grow<1>, grid, dim3(64), 0, st, a, w);
  else
    hipLaunchKernelGGL(k_chain_grow<kChainPlace>, grid, dim3(64), 0, st, a, w);
}
void launch_chain_trials(const ChainArgs &a, const ChainWork &w, const ChainStep &h, hipStream_t st) {
  const int64_t waves = a.n * a.k;
  const dim3 grid((unsigned)((waves + 3) / 4));
  const bool one = one_bead(h);
  if (a.kt <= 128) {
    if (one)
      hipLaunchKernelGGL((k_chain_trials<2, 1>), grid, dim3(256), 0, st, a, w);
    else
      hipLaunchKernelGGL((k_chain_trials<2, kChainPlace>), grid, dim3(256), 0, st, a, w);
  } else {
    if (one)
      hipLaunchKernelGGL((k_chain_trials<kChainTorsionTrials / 64, 1>), grid, dim3(256), 0, st, a, w);
    else
      hipLaunchKernelGGL((k_chain_trials<kChainTorsionTrials / 64, kChainPlace>), grid, dim3(256), 0, st, a, w);
  }
}
void launch_chain_select(const ChainArgs &a, const ChainWork &w, hipStream_t st) {
  hipLaunchKernelGGL(k_chain_select, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a, w);
}
void launch_chain_finalize(const ChainArgs &a, const ChainWork &w, hipStream_t st) {
  hipLaunchKernelGGL(k_chain_finalize, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, st, a, w);
}

// ====================================================================
// Pore-size distribution (src/sample.c:3376-3560), FP64 in the
// reference's operation order so that the counts equal the oracle's.
// ====================================================================
namespace psd {

// ConvertFromABCtoXYZ of three uniforms
__device__ __forceinline__ void point(const Cell &c, uint64_t key, uint64_t ctr, double p[3]) {
  abc_to_xyz(c, stream_uniform(key, ctr), stream_uniform(key, ctr + 1), stream_uniform(key, ctr + 2), p[0], p[1],
             p[2]);
}

// ApplyBoundaryCondition specialised by boundary: the rectangular form keeps
// six cell constants live through the atom loop (the same operations as
// min_image_exact)
template <int BC>
struct Image {
  Cell c;
  double bx, by, bz, ix, iy, iz;
  __device__ explicit Image(const Cell &cell) : c(cell) {
    bx = cell.box[AX];
    by = cell.box[BY];
    bz = cell.box[CZ];
    ix = cell.inv[AX];
    iy = cell.inv[BY];
    iz = cell.inv[CZ];
  }
  __device__ __forceinline__ void apply(double &dx, double &dy, double &dz) const {
#pragma clang fp contract(off)
    if (BC == 1) {
      dx -= bx * nint_d(dx * ix);
      dy -= by * nint_d(dy * iy);
      dz -= bz * nint_d(dz * iz);
    } else
      min_image_exact(c, dx, dy, dz);
  }
};

// CheckSphereOverlap: true when blocked or within 0.5 factor sigma of an
// atom; otherwise radius = min_j (|pos - r_j| - 0.5 factor sigma_j).  The
// atom records are wave-uniform (scalar loads).
template <int BC, int POCK>
__device__ bool overlap(const PsdArgs &a, const Image<BC> &im, const double p[3], double &radius) {
#pragma clang fp contract(off)
  radius = DBL_MAX;
  if (POCK && blocked_pocket(a.pockets, p[0], p[1], p[2])) return true;
  bool ov = false;
  for (int j = 0; j < a.n; ++j) {
    const double4 at = a.atoms[j];
    double dx = p[0] - at.x, dy = p[1] - at.y, dz = p[2] - at.z;
    im.apply(dx, dy, dz);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < at.w * at.w) ov = true;
    // sqrt(rr) - w can undercut the running minimum only when rr < (radius +
    // w)^2; the relative margin keeps every atom within rounding of that
    // bound, so the minimum (and the result) is the one of all atoms
    const double lim = radius + at.w;
    if (rr < lim * lim * (1.0 + 1e-9)) {
      const double r = sqrt(rr) - at.w;
      if (r < radius) radius = r;
    }
    if (__all(ov)) break;
  }
  return ov;
}

// CheckSphereOverlap of two points against the same wave-uniform atom
// records (one scalar load of each record serves both).  The loop ends early
// only when both points of every lane overlap, so a point that does not
// overlap sees every atom: ov and radius equal those of overlap().
template <int BC, int POCK>
__device__ void overlap2(const PsdArgs &a, const Image<BC> &im, const double p[2][3], bool ov[2],
                         double radius[2]) {
#pragma clang fp contract(off)
  bool done[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    radius[u] = DBL_MAX;
    ov[u] = POCK && blocked_pocket(a.pockets, p[u][0], p[u][1], p[u][2]);
    done[u] = ov[u];
  }
  for (int j = 0; j < a.n; ++j) {
    const double4 at = a.atoms[j];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (done[u]) continue;
      double dx = p[u][0] - at.x, dy = p[u][1] - at.y, dz = p[u][2] - at.z;
      im.apply(dx, dy, dz);
      const double rr = dx * dx + dy * dy + dz * dz;
      if (rr < at.w * at.w) ov[u] = true;
      const double lim = radius[u] + at.w;
      if (rr < lim * lim * (1.0 + 1e-9)) {
        const double r = sqrt(rr) - at.w;
        if (r < radius[u]) radius[u] = r;
      }
    }
    if (__all(ov[0] && ov[1])) break;
  }
}

} // namespace psd

template <int BC, int POCK>
__global__ __launch_bounds__(256) void k_psd_a(const PsdArgs a, double *__restrict__ ax) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_cycles) return;
  const uint64_t key = stream_key(a.seed, (uint64_t)(a.first + i));
  const psd::Image<BC> im(a.cell);
  double p[3], radius;
  psd::point(a.cell, key, 0, p);
  const bool ov = psd::overlap<BC, POCK>(a, im, p, radius);
  ax[4 * i + 0] = p[0];
  ax[4 * i + 1] = p[1];
  ax[4 * i + 2] = p[2];
  ax[4 * i + 3] = ov ? 1.0 : 0.0;
}

// blockIdx.y = cycle, blockIdx.x = block of centres; partial[cycle][block]
// = the largest radius of the block's centres that hold A (-10 if none)
template <int BC, int POCK>
__global__ __launch_bounds__(256) void k_psd_b(const PsdArgs a, const double *__restrict__ ax,
                                               double *__restrict__ partial) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.y;
  const int64_t out = i * a.blocks_per_cycle + blockIdx.x;
  if (ax[4 * i + 3] != 0.0) {
    if (threadIdx.x == 0) partial[out] = -10.0;
    return;
  }
  const double A[3] = {ax[4 * i + 0], ax[4 * i + 1], ax[4 * i + 2]};
  const uint64_t key = stream_key(a.seed, (uint64_t)(a.first + i));
  const psd::Image<BC> im(a.cell);
  double largest = -10.0;
  const int64_t j0 = (int64_t)blockIdx.x * a.b_per_block, j1 = j0 + a.b_per_block;
  // two centres per lane (j and j + blockDim.x); every lane runs the atom
  // loop (wave-uniform trip count), out-of-range centres are discarded
  for (int64_t j = j0 + threadIdx.x; j < j1; j += 2 * (int64_t)blockDim.x) {
    const int64_t jj[2] = {j, j + blockDim.x};
    bool live[2];
    double B[2][3];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      live[u] = jj[u] < j1 && jj[u] < a.n_b;
      psd::point(a.cell, key, (uint64_t)(3 + 3 * (live[u] ? jj[u] : 0)), B[u]);
    }
    bool ov[2];
    double radius[2];
    psd::overlap2<BC, POCK>(a, im, B, ov, radius);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (!live[u] || ov[u]) continue;
      double dx = A[0] - B[u][0], dy = A[1] - B[u][1], dz = A[2] - B[u][2];
      im.apply(dx, dy, dz);
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r > radius[u]) continue;
      if (radius[u] > largest) largest = radius[u];
    }
  }
  for (int s = 32; s > 0; s >>= 1) largest = fmax(largest, __shfl_xor(largest, s));
  __shared__ double wmax[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) wmax[wv] = largest;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, wmax[w]);
    partial[out] = m;
  }
}

__global__ __launch_bounds__(256) void k_psd_bin(const PsdArgs a, const double *__restrict__ partial,
                                                 double *__restrict__ largest, unsigned long long *__restrict__ counts) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_cycles) return;
  double m = -10.0;
  for (int b = 0; b < a.blocks_per_cycle; ++b) m = fmax(m, partial[i * a.blocks_per_cycle + b]);
  largest[i] = m;
  const int index = (int)(m / a.delta);
  if ((index >= 0) && (index < a.bins)) atomicAdd(&counts[index], 1ULL);
}

template <int BC, int POCK>
static void psd_launch(const PsdArgs &a, double *ax, double *partial, hipStream_t st) {
  hipLaunchKernelGGL((k_psd_a<BC, POCK>), dim3((unsigned)((a.n_cycles + 255) / 256)), dim3(256), 0, st, a, ax);
  hipLaunchKernelGGL((k_psd_b<BC, POCK>), dim3((unsigned)a.blocks_per_cycle, (unsigned)a.n_cycles), dim3(256), 0,
                     st, a, (const double *)ax, partial);
}

void launch_psd(const PsdArgs &a, double *ax, double *partial, double *largest, unsigned long long *counts,
                hipStream_t st) {
  if (a.n_cycles <= 0) return;
  const bool rect = a.cell.bc == 1, pock = a.pockets.n > 0;
  if (rect && !pock)
    psd_launch<1, 0>(a, ax, partial, st);
  else if (rect)
    psd_launch<1, 1>(a, ax, partial, st);
  else if (!pock)
    psd_launch<2, 0>(a, ax, partial, st);
  else
    psd_launch<2, 1>(a, ax, partial, st);
  hipLaunchKernelGGL(k_psd_bin, dim3((unsigned)((a.n_cycles + 255) / 256)), dim3(256), 0, st, a,
                     (const double *)partial, largest, counts);
}
} // namespace rhg
