// rhg_kernels.h -- launch interface between the host runtime and the gfx950
// kernels of librhg.  All launches are asynchronous on the given stream.
#pragma once

#include "rhg_common.h"

namespace rhg {

// A list of sites (trial bead positions) of one pseudo-atom type.
struct SiteList {
  const double *x, *y, *z; // coordinates, SoA
  int64_t n;               // number of sites (upper bound when n_dev is set)
  const int32_t *n_dev;    // optional device-side count (<= n)
  int32_t count_mult;      // sites = (*n_dev) * count_mult when n_dev is set
  // row layout: with in_stride > 0 the sites are count_mult rows of *n_dev
  // sites each, row r starting at r * in_stride (so that a per-insertion
  // writer stores coalesced); site i of the list is element
  // (i % count_mult) * in_stride + i / count_mult of x, y, z and of the outputs
  int64_t in_stride;
};

struct EnergyOut {
  double *u_vdw; // nullable
  double *u_cc;  // nullable
  uint8_t *overlap; // nullable
};

struct EnergyArgs {
  Cell cell;
  double rc2_vdw, rc2_cc;
  double alpha;  // Ewald alpha
  double cqa;    // C * scaling * q_probe (0 disables the Coulomb term)
  int n_atoms;
  const AtomRec64 *a64; // FP64 records of this probe type
  const AtomRecMx *amx; // mixed records of this probe type
  const float4 *par;    // per atom {4 eps sigma^12, 4 eps sigma^6, shift, q} of this probe type (FP32)
  int precision;        // 0 FP64, 1 mixed
  int do_vdw, do_cc;
  // shifted LJ on the image lists: per cell, the FP64 sum of this type's
  // shifts over the cell's list (launch_img_shift_sums); null: not used
  const double *cell_shift = nullptr;
};

void launch_site_energy(const EnergyArgs &a, const SiteList &s, const EnergyOut &o,
                        hipStream_t st);

// ------------------------------------------------------------ cell lists
// Cells are a regular nc[0] x nc[1] x nc[2] split of fractional space.  The
// neighbour list of a cell holds, in ascending order, every atom whose
// periodic (torus) distance to the cell centre is at most rc + R_cell, so it
// contains every atom within rc of any point of the cell.
struct CellGrid {
  int nc[3];
  int ncell;
  const int32_t *nb_off; // ncell + 1
  const int32_t *nb_idx; // concatenated ascending atom indices
  // image lists (nullable): per cell every periodic image of every atom
  // within reach of the cell centre, {dx, dy, dz} in A as FP32 and the atom
  // index bit-cast into w
  const int32_t *img_off; // ncell + 1
  const float4 *img;
};

// sites per lane of the image-list kernels: their tiles hold 64 kImgLane
// sites of one cell
constexpr int kImgLane = 4;

constexpr int kPairLanes = 256; // counters per kind of the executed-pair tallies

// scratch for binning one site list into cells
struct BinWork {
  int32_t *cell_of;  // per site
  int32_t *count;    // per cell (cleared by the scan once read)
  int32_t *rank;     // per site: its place among its cell's sites
  int32_t *start;    // per cell, exclusive scan of count
  int32_t *tile_off; // per cell + 1, exclusive scan of ceil(count/ts): ts = 256
                     // sites when fix is set (image-list kernels), else 64
  int32_t *perm;     // sorted position -> site index (written only when fix is not)
  // optional (image-list kernels): per sorted position the site's offset from
  // its cell centre in A as FP32 and the site index bit-cast into w
  float4 *fix;
  // optional, with fix: per site in list order its FP32 offset from its cell
  // centre and the cell index bit-cast into w, written by the counting pass
  // (in place of cell_of), so that the scatter moves 16 B per site and reads
  // no positions
  float4 *fixu = nullptr;
  // image-list kernels: (sorted position, cell) of the sites to recompute
  // exactly (room for every site) and their count (cleared by the binning)
  int2 *redo = nullptr;
  int32_t *n_redo = nullptr;
  // executed pair evaluations by kind (LJ, LJ+Coulomb, Coulomb), each kind
  // spread over kPairLanes counters (kind * kPairLanes + block % kPairLanes)
  // so that the per-wave atomics do not serialise on one address
  unsigned long long *pairs;
  // launch shape of the site kernels: sites per tile of the image-list
  // kernels (64, 128, 192 or 256; 0: 256; the kernel's registers are sized
  // for it) and waves per workgroup (1, 2 or 4; 0: 2)
  int32_t tile_sites = 0;
  int32_t waves = 0;
  // the counts are known to be zero (fresh scratch cleared at allocation, or
  // left cleared by the previous binning's scan): no clearing launch
  int32_t counts_zero = 0;
};

__host__ __device__ inline int img_tile_sites(const BinWork &b) { return b.tile_sites ? b.tile_sites : 64 * kImgLane; }
inline int site_waves(const BinWork &b) { return b.waves ? b.waves : 2; }

void launch_cell_neighbours(const Cell &cell, int nc0, int nc1, int nc2, double reach,
                            const double *ax, const double *ay, const double *az, int n_atoms,
                            int32_t *counts, int32_t *nb_off, int32_t *nb_idx, int pass,
                            hipStream_t st);
void launch_scan_counts(const int32_t *counts, int n, int32_t *off, hipStream_t st);
// pass 0 counts the images of each cell's neighbour atoms into counts, pass 1
// writes them at off
// Neighbour lists and image lists of every cell in one walk over the atoms,
// entry for entry those of launch_cell_neighbours followed by
// launch_image_list: pass 0 counts both into nb_counts and img_counts,
// pass 1 writes them at nb_off and img_off
void launch_cell_lists(const Cell &cell, const CellGrid &g, double reach, const double widths[3],
                       double box_reach, const double *ax, const double *ay, const double *az, int n_atoms,
                       int32_t *nb_counts, int32_t *img_counts, const int32_t *nb_off, int32_t *nb_idx,
                       const int32_t *img_off, float4 *img, int pass, hipStream_t st);
// exclusive scans of in0 and in1 (n counts each) into off0 and off1 (n + 1),
// tot[0] and tot[1] the totals
void launch_scan2(const int32_t *in0, int32_t *off0, const int32_t *in1, int32_t *off1, int n, int32_t *tot,
                  hipStream_t st);
// widths: perpendicular widths of the cell
// out[c] = sum of a64[j].shift over the image list of cell c, c < g.ncell
void launch_img_shift_sums(const CellGrid &g, const AtomRec64 *a64, double *out, hipStream_t st);
// box_reach > 0: on rectangular cells keep the images within box_reach of the
// cell's box (the cut-off plus a rounding margin) rather than within reach
// of its centre
void launch_image_list(const Cell &cell, const CellGrid &g, double reach, const double widths[3],
                       double box_reach, const double *ax, const double *ay, const double *az,
                       int32_t *counts, const int32_t *off, float4 *img, int pass, hipStream_t st);
void launch_bin_sites(const Cell &cell, const CellGrid &g, const SiteList &s, const BinWork &b,
                      hipStream_t st);
// Start and stop events that the dispatch of a site-energy kernel records
// itself (hipExtLaunchKernelGGL): the kernel's own time, with no marker
// packets between launches; both null: a plain launch.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
void launch_site_energy_culled(const EnergyArgs &a, const CellGrid &g, const SiteList &s,
                               const BinWork &b, const EnergyOut &o, hipStream_t st,
                               LaunchEvents ev = LaunchEvents());
// n interleaved positions (x, y, z per site) into three SoA arrays
void launch_deinterleave(const double *xyz, int64_t n, double *x, double *y, double *z, hipStream_t st);

// ---------------------------------------------------------------- Ewald
// largest k_x of the Fourier kernel (k_x + 1 slices)
constexpr int kMaxKx = 64;

struct EwaldDev {
  int kx, ky, kz, nk;
  double alpha, uion, net_q_fw, beta_unused;
  const double *kfactor;   // nk, RASPA2 order
  const double2 *sf;       // nk framework structure factors
  const int32_t *row_off;  // (kx+1)*(2ky+1): nvec index of the row's first valid kz
  const int32_t *row_kz;   // (kx+1)*(2ky+1): largest valid |kz| of the row (-1 empty)
  // the k-vectors in the Fourier kernel's visiting order, {K Re S_fw,
  // K Im S_fw, K, index} (nk records and 8 of padding), and the first
  // record of each kx slice (kst_off[kx + 1] = nk)
  const double4 *kst;
  int kst_off[kMaxKx + 2];
  Cell cell;
  double C;                // Coulomb conversion factor
};

// structure factors of the charged framework atoms (fractional coordinates)
// tab: scratch of n x ((kx+1) + (2ky+1) + (2kz+1)) double2 (phase powers)
void launch_structure_factor(const EwaldDev &e, const double *fx, const double *fy, const double *fz,
                             const double *q, int n, double2 *tab, double2 *out, hipStream_t st);

// kst[i] = {kf[k] sf[k].x, kf[k] sf[k].y, kf[k], k} with k = idx[i], i < n
void launch_kstream(const double *kf, const double2 *sf, const int32_t *idx, int n, double4 *kst,
                    hipStream_t st);

// Fourier Delta of n molecules; positions pos[j][m] (site-major SoA of
// three arrays each n_sites*stride); list/n_dev select molecules.
struct FourierArgs {
  EwaldDev e;
  int n_sites;
  int nc;                 // charged sites
  int cidx[kMaxSites];    // site index of each charged site, in site order
  double q[kMaxSites];    // charge of each charged site (negated for a removed molecule's)
  int grp[kMaxSites];     // 0: new molecule, 1: removed molecule (excluded pairs per group)
  const double2 *sads;    // adsorbate structure factor (StoreTotalChargeAdsorbates), nullable
  double net_q_ads;       // NetChargeAdsorbates
  double self_energy; // sum C q^2 alpha/sqrt(pi) of the charged sites
  const double *px, *py, *pz; // [site * stride + m]
  int64_t stride;
  int64_t n;                  // molecules (upper bound with n_dev)
  const int32_t *n_dev;
  double *d_host, *d_ads;     // per molecule
};
// part: (kx+1) x stride double2 scratch for the per-kx partial sums
void launch_fourier(const FourierArgs &a, double2 *part, hipStream_t st);

// ------------------------------------------------------- finite loading
// Guest-guest terms (src/inter_energy.c:321-520): the adsorbate atoms as
// seen by one probe pseudo-atom, molecule index per atom, FP64
struct AdsArgs {
  Cell cell;
  const AtomRec64 *rec; // n records of this probe type
  const int32_t *mol;   // molecule of each atom
  int n;
  int exclude;          // molecule left out (-1 or the number of molecules: none)
  // per-insertion exclusions (nullable): the site list's column (physical
  // index modulo its row stride) is an insertion, or through col_map (nullable)
  // the slot of one
  const int32_t *excl;
  const int32_t *col_map;
  double rc2v, rc2c, alpha, cqa;
  int do_vdw, do_cc;
};
void launch_site_ads(const AdsArgs &a, const SiteList &s, double *u_vdw, double *u_cc, hipStream_t st);

// ---------------------------------------------------------------- Widom
struct WidomWork {
  int64_t cap;       // insertions per chunk
  // first-bead trials, stored [trial][insertion] (row stride cap)
  double *fbx, *fby, *fbz, *fb_uv, *fb_uc;
  // per insertion state
  int32_t *stage, *selfb, *selor;
  uint64_t *fbmask, *ormask;
  double *W1, *W2, *W, *dU, *uvdw, *ureal, *ufh, *ufa, *usel_or;
  double *selx, *sely, *selz; // selected first-bead position
  // compaction after the first-bead stage
  int32_t *active, *n_active;
  // orientation trials of the placed beads, stored
  // [(storage slot of the bead) * k_or + orientation][active slot]
  double *orx, *ory, *orz, *or_uv;
  // positions of the charged placed beads of the selected orientations,
  // [charged storage slot][slot2], grouped by pseudo-atom type
  double *ccx, *ccy, *ccz;
  // compaction after the orientation stage; selected configuration [j][slot2]
  int32_t *active2, *n_active2;
  double *cfx, *cfy, *cfz;
  double *cc_place; // [charged placed slot][slot2] real-space Coulomb of the placed beads
  double *f_host, *f_ads; // [slot2]
  // guest-guest terms at finite loading (all null without adsorbates), laid
  // out as fb_uv, or_uv and cc_place; per insertion the selected sums
  double *fb_av, *fb_ac, *or_av, *cc_ads, *uadsv, *uadsc;
  // host-supplied trials (nullable); h_old: [insertion][site][3] positions
  // of retraced molecules
  const double *h_fb, *h_cord, *h_usel, *h_old;
};

// Control block of a Widom run, device-resident so that a captured chunk
// can be replayed: ctl[0] = job index of the chunk's first insertion,
// ctl[1] = end of the submitted range, ctl[2..7] = job indices where the
// five blocks start (and the job end).
enum { CTL_FIRST = 0, CTL_END = 1, CTL_JB = 2, CTL_N = 8 };

// Blocked pockets (src/framework.c:6168-6300): centres periodic with the
// unit cell; a grown molecule with an atom inside one (outside all, when
// inverted) is rejected (src/mc_moves.c:7882-7886)
struct PocketArgs {
  int n;            // pockets (0: none)
  int invert;
  int bc;           // 1 rectangular, otherwise triclinic
  const double *pk; // n x 4: x, y, z, radius
  double ubox[9], uinv[9];
};

// ------------------------------------------------------- surface area
// SurfaceAreaMove / CheckSurfaceAreaOverlap (src/mc_moves.c:16434-16520,
// src/framework.c:11649-11720): per framework atom i, points on the sphere
// of radius d[i] = factor * sigma(probe, i) around it; a point counts when it
// is in no blocked pocket and within d[j] of no other atom j.
struct SurfaceArgs {
  Cell cell;
  PocketArgs pockets;
  const double4 *atoms; // x, y, z, d per atom
  int n;
  int64_t samples;      // points per atom
  uint64_t seed;
};
// counts[i] += accessible points of atom i (counts zeroed by the caller)
void launch_surface_area(const SurfaceArgs &a, unsigned long long *counts, hipStream_t st);

struct WidomArgs {
  PocketArgs pockets;
  MolDev mol;
  Cell cell;
  double beta;
  uint64_t seed;
  int64_t n;           // capacity of a chunk (grid size); the chunk runs min(n, end - first)
  const int64_t *ctl;  // control block (above)
  int32_t *n_eff;      // device: insertions of this chunk, written by fb_gen
  int use_host_trials;
  int ewald;
  // retrace of existing molecules (RetraceMolecule, CBMC_DELETION): trial 0
  // of both stages is the molecule's configuration (h_old) and is always
  // chosen; no minimum-Rosenbluth, blocked-pocket or Fourier step
  int retrace;
  int exclude; // adsorbate molecule left out of the guest-guest terms
  const int32_t *excl; // per insertion instead of `exclude` (device, nullable)
};

__device__ __forceinline__ int64_t chunk_first(const WidomArgs &a) { return a.ctl[CTL_FIRST]; }
__device__ __forceinline__ int64_t chunk_n(const WidomArgs &a) {
  const int64_t left = a.ctl[CTL_END] - a.ctl[CTL_FIRST];
  return left < a.n ? (left > 0 ? left : 0) : a.n;
}

// int BlockedPocket(VECTOR) for n positions (xyz interleaved) -> flags
void launch_blocked_pockets(const PocketArgs &p, const double *xyz, int64_t n, uint8_t *flags,
                            hipStream_t st);

// writes the control block (values by argument, no host memory involved)
void launch_set_ctl(int64_t *ctl, int64_t first, int64_t end, const int64_t jb[6], hipStream_t st);
void launch_widom_fb_gen(const WidomArgs &a, const WidomWork &w, hipStream_t st);
void launch_widom_fb_select(const WidomArgs &a, const WidomWork &w, hipStream_t st);
void launch_widom_or_select(const WidomArgs &a, const WidomWork &w, hipStream_t st);
void launch_widom_finalize(const WidomArgs &a, const WidomWork &w, hipStream_t st);
void launch_widom_records(const WidomArgs &a, const WidomWork &w, double *rec, hipStream_t st);
void launch_zero_counters(int32_t *a, int32_t *b, hipStream_t st);

// deterministic block sums over insertions [0,n) of a chunk; bounds[6] are
// the chunk-local insertion index where each block starts (clamped).
struct BlockPartials {
  double *part; // n_seg x 5 x 3
  int n_seg;
};
// Adds the chunk's per-block sums of W and W*dU and insertion counts to
// acc[15] (sum_w[5], sum_wdu[5], count[5]) in a fixed order; jb[6] are the
// job-global insertion indices where each block starts.
void launch_block_reduce(const WidomArgs &a, const double *W, const double *dU, double *part, int seg,
                         double *acc, hipStream_t st);

// Adds the chunk's work counts to stats[9] (see rhg_widom_stats).
struct StatArgs {
  double n_atoms, nk;
  int k_fb, k_or, n_place, n_charged, n_cc_place, fb_charged, ewald, multi;
};
void launch_widom_stats(const WidomArgs &a, const WidomWork &w, const StatArgs &sa, double *stats,
                        hipStream_t st);
// ctl[CTL_FIRST] += step (moves a captured chunk on to the next insertions)
void launch_advance(int64_t *ctl, int64_t step, hipStream_t st);
// the block reduction, the work counters and the advance of a chunk in two
// launches (k_block_part, then k_chunk_tail in one workgroup)
void launch_chunk_tail(const WidomArgs &a, const WidomWork &w, const StatArgs &sa, const double *W,
                       const double *dU, double *part, int seg, double *acc, double *stats, int64_t *ctl,
                       int64_t step, hipStream_t st);

// streaming copy of `bytes` (a multiple of 16) for byte-counter calibration
void launch_stream_copy(const void *src, void *dst, int64_t bytes, hipStream_t st);

// ----------------------------------------------------------------- grids
struct GridArgs {
  Cell cell;
  int n[3];          // NumberOfVDWGridPoints (nodes per axis = n+1)
  double size[3], shift[3], delta[3];
  double rc2;
  int n_atoms;
  const AtomRec64 *a64;
  const double *c24, *c96, *c384; // per atom: 24*eps, 96*eps, 384*eps
  int fp64;
  void *out;         // [i][j][k][8] float or double
  int64_t node0, n_nodes; // node range of this launch
  // Ewald-Coulomb grid of a unit charge (coulomb = 1): atoms as {x, y, z, q},
  // rc2 is the charge cut-off squared
  int coulomb;
  const double4 *aq;
  double cc, alpha, sqrtpi; // conversion factor C, Ewald alpha, sqrt(pi)
};
void launch_grid_build(const GridArgs &g, hipStream_t st);
// node positions of nodes [node0, node0 + n_nodes) into x, y, z
void launch_grid_nodes(const GridArgs &g, double *x, double *y, double *z, hipStream_t st);
// culled build of the same nodes, binned as site list s (site = node - node0):
// precision 0 walks the neighbour lists in atom order (bit-identical to
// launch_grid_build), 1 the image lists with FP32 terms
void launch_grid_culled(const GridArgs &g, const CellGrid &cg, const SiteList &s, const BinWork &b,
                        const float4 *par, int precision, hipStream_t st);
void launch_derivatives(const GridArgs &g, const double *xyz, int64_t n, double *out,
                        hipStream_t st);
// CalculateDerivativesAtPositionReal at n positions: out[n][14] as
// launch_derivatives, smallest[n] the smallest in-range distance
void launch_derivatives_real(const GridArgs &g, const double *xyz, int64_t n, double *out,
                             double *smallest, hipStream_t st);
// node-major [node][8] <-> channel-major [8][node] float channels (the file
// layout of grids); to_file 1 gathers node-major into channel-major
void launch_grid_transpose(const float *src, float *dst, int64_t nodes, int to_file, hipStream_t st);

struct InterpArgs {
  int n[3];
  double size[3], shift[3];
  double ubox[9], uinv[9];
  double ulen[3];
  int bc;
  int fp64;
  int coulomb;   // no overlap short-cut; the result is scaled by charge
  double charge;
  const void *grid;
};
void launch_interp(const InterpArgs &a, const double *xyz, int64_t n, double *out,
                   hipStream_t st);
// site energies from grids (gv and/or gc non-null); s.n bounds the grid of
// threads, s.n_dev (when set) the live sites
void launch_site_grid(const InterpArgs *gv, const InterpArgs *gc, const SiteList &s, const EnergyOut &o,
                      int merge, hipStream_t st);

// ------------------------------------------------ pore-size distribution
// SamplePoreSizeDistribution / CheckSphereOverlap (src/sample.c:3376-3560):
// per cycle a point A and n_b sphere centres B from the cycle's counter
// stream; the largest free sphere (radius = distance to the nearest atom
// minus half its scaled sigma) whose centre holds A.
struct PsdArgs {
  Cell cell;
  PocketArgs pockets;
  const double4 *atoms; // x, y, z, 0.5 factor sigma
  int n;
  uint64_t seed;
  int64_t first, n_cycles, n_b;
  int b_per_block;      // centres per block (blockDim.x x per-thread count)
  int blocks_per_cycle;
  int bins;
  double delta;         // PSDRange / PSDHistogramSize
};
// largest[i] per cycle (-10 when none) via partial[i][blocks_per_cycle],
// then counts[index] += 1 for the cycles whose index is in range
void launch_psd(const PsdArgs &a, double *ax, double *partial, double *largest,
                unsigned long long *counts, hipStream_t st);

// ------------------------------------------------------ flexible chains
// CBMC growth of a flexible molecule of one group (src/cbmc.c:2762-2849,
// 1480-2482, 1154-1478).  The growth path is the same for every insertion;
// the host resolves it into ChainStep tables (device memory, read
// wave-uniformly).  Atom references inside a step are local codes: 0..3 the
// beads placed by the step, kLocCur the current bead, kLocPrev the previous
// one, kLocGlobal + atom any other placed atom.
constexpr int kChainMax = 64;   // atoms of a chain
constexpr int kChainPlace = 4;  // beads placed by one step
constexpr int kChainBonds = 4, kChainBends = 8, kChainTors = 16, kChainIntra = 64;
constexpr int kChainTorsionTrials = 1024;
constexpr int kLocCur = 4, kLocPrev = 5, kLocGlobal = 8;
// per-insertion chain record fields (the C ABI's RHG_CREC_*)
enum {
  kCrecW = 0, kCrecDU, kCrecUHost, kCrecUIntra, kCrecUBond, kCrecUBend, kCrecUTors, kCrecOverlap,
  kCrecStage, kCrecStep, kCrecW1, kCrecN
};

struct ChainStep {
  int cur, prev, n_prev;
  int n_place;                  // beads placed (= branches when n_prev = 1)
  int place[kChainPlace];       // atom ids
  int type[kChainPlace];        // pseudo-atom of each placed bead
  int n_bonds, n_bends, n_tors, n_intra;
  int bond_a[kChainBonds], bond_b[kChainBonds], bond_kind[kChainBonds]; // local codes; 0 fixed, 1 harmonic
  int bond_branch[kChainBonds];  // branch slot whose bond this is (BranchAtoms[jj][0])
  double bond_p0[kChainBonds], bond_p1[kChainBonds]; // p0 / ENERGY_TO_KELVIN, p1 [A] (fixed: p0 = length)
  int bend_a[kChainBends], bend_b[kChainBends], bend_c[kChainBends];
  double bend_p0[kChainBends], bend_p1[kChainBends]; // internal units, radians
  int tors_a[kChainTors], tors_b[kChainTors], tors_c[kChainTors], tors_d[kChainTors];
  double tors_p[kChainTors][4];
  int intra_a[kChainIntra], intra_b[kChainIntra];
  double intra_scal[kChainIntra], intra_c4[kChainIntra], intra_s2[kChainIntra], intra_shift[kChainIntra];
  double ref_len[kChainPlace];   // GetReferenceBondLength(cur, branch)
  double max_bond[kChainPlace];  // MaximumCBMCChangeBondLength of each branch atom
  double max_bend, max_cone;     // ...ChangeBendAngle / ...RotationOnCone of the current bead
};

struct ChainWork {
  int64_t cap;
  // first-bead trials [trial][insertion] and per-insertion results
  double *fbx, *fby, *fbz, *fb_uv;
  double *W, *dU;             // [insertion]
  double *rec;                // [insertion][RHG_CREC_N] (nullable)
  int32_t *active, *n_active; // slot -> insertion
  // per slot
  int32_t *alive;             // 1 growing, 0 done (rejected)
  double *sW, *uhost, *uintra, *ubond, *ubend, *utors, *W1;
  double *px, *py, *pz;       // grown positions [atom][slot]
  double *bx, *by, *bz;       // internal MC result [place][slot]
  double *bl;                 // bond length of a sphere step [slot]
  // trials of the step: positions [place * k + t][slot], energies [t][slot]
  double *tx, *ty, *tz, *tuv;
  double *tRT, *tbond, *tbend, *ttors, *tintra;
  double *xyz;                // [insertion][atom][3] grown molecule (nullable)
};

struct ChainArgs {
  Cell cell;
  PocketArgs pockets;
  double beta;
  uint64_t seed;
  int64_t first, n;     // insertions [first, first + n) of the chunk
  double rc2;           // CutOffVDW^2 (intra-molecular pairs)
  int n_atoms, start, k_fb, k, kt, moves;
  int step;             // 1-based growth step
  const ChainStep *st;  // device table of this step
};

void launch_chain_fb_select(const ChainArgs &a, const ChainWork &w, hipStream_t st);
void launch_chain_grow(const ChainArgs &a, const ChainWork &w, const ChainStep &h, hipStream_t st);
void launch_chain_trials(const ChainArgs &a, const ChainWork &w, const ChainStep &h, hipStream_t st);
void launch_chain_select(const ChainArgs &a, const ChainWork &w, hipStream_t st);
void launch_chain_finalize(const ChainArgs &a, const ChainWork &w, hipStream_t st);

} // namespace rhg
