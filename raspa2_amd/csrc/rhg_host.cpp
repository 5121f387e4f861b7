// rhg_host.cpp -- host runtime and C ABI of librhg (see include/rhg.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rhg.h"
#include "rhg_internal.h"
#include "rhg_kernels.h"

using namespace rhg;

namespace {

// src/constants.h:48-60 and src/constants.c:141-180
constexpr double kAngstrom = 1e-10, kPicoSecond = 1e-12, kAmu = 1.6605402e-27,
                 kElectron = 1.60217733e-19, kMolarGas = 8.314464919,
                 kBoltzmann = 1.380650324e-23, kAvogadro = 6.0221419947e23,
                 kEpsilon0 = 8.8541878176e-12;

struct Units {
  double E, C, E2K, K2E, KB;
};
Units make_units() {
  volatile double length = kAngstrom, time = kPicoSecond, mass = kAmu, charge = kElectron,
                  dielectric = 1.0;
  Units u;
  u.E = ((mass) * (length * length) / (time * time));
  u.KB = kBoltzmann / u.E;
  u.C = (((charge * charge) / (4.0 * M_PI * kEpsilon0 * length * u.E))) / (dielectric);
  u.E2K = ((u.E * kAvogadro) / kMolarGas);
  u.K2E = (kMolarGas / (u.E * kAvogadro));
  return u;
}

inline int nint(double x) { return (int)((x) >= 0.0 ? ((x) + 0.5) : ((x)-0.5)); }

// Invert3x3Matrix (src/matrix.c:902-927)
void invert3x3(const double *in, double *out) {
  out[AX] = in[BY] * in[CZ] - in[BZ] * in[CY];
  out[AY] = in[AZ] * in[CY] - in[AY] * in[CZ];
  out[AZ] = in[AY] * in[BZ] - in[AZ] * in[BY];
  out[BX] = in[BZ] * in[CX] - in[BX] * in[CZ];
  out[BY] = in[AX] * in[CZ] - in[AZ] * in[CX];
  out[BZ] = in[AZ] * in[BX] - in[AX] * in[BZ];
  out[CX] = in[BX] * in[CY] - in[BY] * in[CX];
  out[CY] = in[AY] * in[CX] - in[AX] * in[CY];
  out[CZ] = in[AX] * in[BY] - in[AY] * in[BX];
  double d = in[AX] * out[AX] + in[BX] * out[AY] + in[CX] * out[AZ];
  double r = 0.0;
  if (fabs(d) > 0.0) r = 1.0 / d;
  for (int i = 0; i < 9; ++i) out[i] *= r;
}

// CellProperties (src/framework.c:6363-6391): lengths, perpendicular widths, volume
void cell_properties(const double *in, double *p) {
  p[0] = sqrt(in[AX] * in[AX] + in[AY] * in[AY] + in[AZ] * in[AZ]);
  p[1] = sqrt(in[BX] * in[BX] + in[BY] * in[BY] + in[BZ] * in[BZ]);
  p[2] = sqrt(in[CX] * in[CX] + in[CY] * in[CY] + in[CZ] * in[CZ]);
  double axb1 = in[AY] * in[BZ] - in[AZ] * in[BY];
  double axb2 = in[AZ] * in[BX] - in[AX] * in[BZ];
  double axb3 = in[AX] * in[BY] - in[AY] * in[BX];
  double bxc1 = in[BY] * in[CZ] - in[BZ] * in[CY];
  double bxc2 = in[BZ] * in[CX] - in[BX] * in[CZ];
  double bxc3 = in[BX] * in[CY] - in[BY] * in[CX];
  double cxa1 = in[CY] * in[AZ] - in[AY] * in[CZ];
  double cxa2 = in[AX] * in[CZ] - in[AZ] * in[CX];
  double cxa3 = in[AY] * in[CX] - in[AX] * in[CY];
  double vol = fabs(in[AX] * bxc1 + in[AY] * bxc2 + in[AZ] * bxc3);
  p[3] = vol / sqrt(bxc1 * bxc1 + bxc2 * bxc2 + bxc3 * bxc3);
  p[4] = vol / sqrt(cxa1 * cxa1 + cxa2 * cxa2 + cxa3 * cxa3);
  p[5] = vol / sqrt(axb1 * axb1 + axb2 * axb2 + axb3 * axb3);
  p[6] = vol;
}

Cell make_cell(const double *box, int bc) {
  Cell c;
  memcpy(c.box, box, sizeof(double) * 9);
  invert3x3(c.box, c.inv);
  c.bc = bc;
  c.upper = (box[AY] == 0.0 && box[AZ] == 0.0 && box[BZ] == 0.0 && c.inv[AY] == 0.0 &&
             c.inv[AZ] == 0.0 && c.inv[BZ] == 0.0)
                ? 1
                : 0;
  return c;
}

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  // buffers outgrown while work may still be queued on them; freed with the
  // owner (hipFree waits for the whole device, which would stall the other
  // streams of a server)
  std::vector<std::pair<void *, size_t>> retired;
};

struct PseudoAtom {
  double eps_k, sigma, q;
  AtomRec64 *d64 = nullptr;
  AtomRecMx *dmx = nullptr;
  double *c24 = nullptr, *c96 = nullptr, *c384 = nullptr;
  float4 *par = nullptr; // {4 eps sigma^12, 4 eps sigma^6, shift, q} for the image-list kernels
  AtomRec64 *dads = nullptr; // the adsorbate atoms as seen by this type (finite loading)
  size_t dads_cap = 0;
  // per cell of the image lists: the sum of this type's LJ shifts over the
  // cell's list, for image lists of version cell_shift_version
  double *cell_shift = nullptr;
  int cell_shift_cap = 0;
  uint64_t cell_shift_version = 0;
};

struct Pending {
  std::string name;
  std::string stage; // also recorded under this family when not empty
  hipEvent_t a, b;
};

// Device memory of a context's fixed tables: small buffers are carved out of
// 2 MiB blocks, large ones get a block each, so that setting up a framework
// costs a handful of hipMalloc calls (they contend across the threads of a
// screening server) instead of one per table.  Carved buffers are released
// with the context; large ones can be released early.
struct Arena {
  std::vector<std::pair<char *, size_t>> small; // carved blocks
  std::vector<std::pair<char *, size_t>> large; // one buffer each
  size_t used = 0;                              // bytes carved from the last small block
  rhg_pool *pool = nullptr;                     // where blocks come from (null: hipMalloc)
};

} // namespace

// Device memory shared by the contexts of a screening server: slabs taken
// with hipMalloc (one call per slab instead of one per table: hipMalloc and
// hipFree serialise across the server's threads), carved at the tail or
// reused from blocks that evicted contexts gave back (smallest block that
// fits; a remainder of 2 MiB or more goes back to the free blocks).
struct rhg_pool {
  int device = 0;
  size_t slab = 0;
  std::mutex mu;
  std::vector<std::pair<char *, size_t>> slabs;
  size_t tail = 0; // bytes carved from the last slab
  std::multimap<size_t, char *> free_blocks;
};

rhg_pool *rhg_pool_create(int device, size_t slab_bytes) {
  rhg_pool *p = new rhg_pool();
  p->device = device;
  p->slab = std::max<size_t>(slab_bytes, (size_t)64 << 20);
  return p;
}

void rhg_pool_destroy(rhg_pool *p) {
  if (!p) return;
  hipSetDevice(p->device);
  for (auto &sl : p->slabs) hipFree(sl.first);
  delete p;
}

namespace {

constexpr size_t kPoolSplit = (size_t)2 << 20;

hipError_t pool_alloc(rhg_pool *P, void **out, size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> lk(P->mu);
  auto it = P->free_blocks.lower_bound(bytes);
  if (it != P->free_blocks.end()) {
    const size_t have = it->first;
    char *b = it->second;
    P->free_blocks.erase(it);
    if (have - bytes >= kPoolSplit) P->free_blocks.emplace(have - bytes, b + bytes);
    *out = b;
    return hipSuccess;
  }
  if (P->slabs.empty() || P->tail + bytes > P->slabs.back().second) {
    if (!P->slabs.empty() && P->slabs.back().second - P->tail >= kPoolSplit)
      P->free_blocks.emplace(P->slabs.back().second - P->tail, P->slabs.back().first + P->tail);
    void *sl = nullptr;
    const size_t n = std::max(P->slab, bytes);
    hipError_t e = hipMalloc(&sl, n);
    if (e != hipSuccess) return e;
    P->slabs.push_back({(char *)sl, n});
    P->tail = 0;
  }
  *out = P->slabs.back().first + P->tail;
  P->tail += bytes;
  return hipSuccess;
}

void pool_free(rhg_pool *P, void *p, size_t bytes) {
  if (!p) return;
  bytes = (bytes + 255) & ~(size_t)255;
  std::lock_guard<std::mutex> lk(P->mu);
  P->free_blocks.emplace(bytes, (char *)p);
}

hipError_t dev_malloc(rhg_pool *P, void **out, size_t bytes) {
  return P ? pool_alloc(P, out, bytes) : hipMalloc(out, bytes);
}

void dev_free(rhg_pool *P, void *p, size_t bytes) {
  if (P)
    pool_free(P, p, bytes);
  else if (p)
    hipFree(p);
}

} // namespace

struct rhg_workspace {
  int device = 0;
  std::map<std::string, DevBuf> bufs;
  uint64_t gen = 0;
  rhg_ctx *owner = nullptr;
  // process-wide unique id: captured graphs are keyed by it, never by the
  // address, which a new workspace may reuse after this one is destroyed
  uint64_t uid = 0;
};

struct rhg_ctx {
  int device = 0;
  rhg_workspace *ws = nullptr; // borrowed Widom/scratch buffers (nullable)
  hipStream_t stream = nullptr;
  bool own_stream = true;      // false: the stream is borrowed (screening server)
  // a second stream of a context that owns its stream (created on first
  // use): a Widom chunk's Fourier stage runs there beside the placed-bead
  // Coulomb stage, joined back before finalisation (default priority)
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  Arena arena;
  std::string err;
  Cell cell;
  double props[7];
  int n = 0;
  std::vector<double> x, y, z, q, eps, sig;
  double mass_total = 0.0;
  double rc_vdw = 12.0, rc2_vdw = 144.0;
  int shifted = 1;
  int charge_method = RHG_CHARGE_NONE;
  double rc_cc = 12.0, rc2_cc = 144.0;
  double temperature = 300.0;
  Units u;
  double beta = 0.0;
  int precision = RHG_PREC_FP64;
  std::vector<PseudoAtom> types;
  // Ewald
  double alpha = 0.0, uion = 0.0, net_q_fw = 0.0, recip_cutoff = 0.0, ewald_precision = 0.0;
  int kx = 0, ky = 0, kz = 0, nk = 0;
  std::vector<double> kfactor;
  std::vector<int32_t> kidx, row_off, row_kz;
  double *d_kfactor = nullptr;
  double2 *d_sf = nullptr;
  int32_t *d_row_off = nullptr, *d_row_kz = nullptr;
  // k-vector records in the Fourier kernel's visiting order and each kx
  // slice's first record
  double4 *d_kst = nullptr;
  std::vector<int32_t> kst_off;
  // the framework structure factor and the Fourier records are formed on the
  // first Fourier use (a request with uncharged probes never needs them)
  const int32_t *d_korder = nullptr;
  int korder_n = 0;
  bool sf_ready = false;
  // host copies of uploads that may still be in flight, released after the
  // next synchronisation of the stream (so that set-up need not wait for them)
  std::vector<std::vector<char>> kept;
  size_t footprint = 0;
  // cell lists
  int cull = 1;
  double cell_size = 2.5;
  // launch shape of the site kernels (rhg_set_tiles; 0: defaults)
  int tile_sites = 0, waves = 0;
  bool grid_ready = false;
  CellGrid cg;
  int32_t *d_nb_off = nullptr, *d_nb_idx = nullptr;
  int32_t *d_img_off = nullptr;
  float4 *d_img = nullptr;
  bool img_ready = false;
  uint64_t img_version = 0; // counts image-list builds
  int use_images = 1; // mixed precision walks the image lists when they exist
  double cell_reach = 0.0, cell_halfdiag = 0.0;
  double *d_ax = nullptr, *d_ay = nullptr, *d_az = nullptr;
  std::map<std::string, DevBuf> bufs;
  // kernel timing with HIP events: 0 off, 1 every kernel family, 2 only the
  // site-energy kernels (RHG_PROFILE_SITE_ENERGY)
  int prof = 0;
  // work counters of the Widom path since the last reset:
  // insertions, first-bead sites, orientation sites, Coulomb sites of the
  // selected orientations, Fourier molecules, LJ pairs, LJ+Coulomb pairs,
  // Coulomb-only pairs, k-vector x charged-site terms
  // device accumulators of the Widom path: [0,15) block sums of the
  // current collection (sum_w[5], sum_wdu[5], count[5]), [16,25) the work
  // counters above since the last reset
  double *d_acc = nullptr;
  int64_t *d_ctl = nullptr;
  int32_t *d_neff = nullptr;
  int64_t chunk = (int64_t)1 << 20;
  int graphs = 1;
  // buffer generation: bumped whenever a device pointer that a captured
  // chunk holds may have changed
  uint64_t gen = 0;
  std::map<std::string, hipGraphExec_t> graph_cache;
  // host-side fault hook: fail the chunk after this many more chunks (<0 off)
  int64_t fault_after = -1;
  // ideal-gas Rosenbluth sums of the insertions enqueued since the last
  // collect, per block: count x V / (N + 1) with N the molecules present
  // when they were enqueued (src/mc_moves.c:7861-7863)
  double ig_pending[RHG_NR_BLOCKS] = {0.0, 0.0, 0.0, 0.0, 0.0};
  // adsorbates present (finite loading): molecule m holds atoms
  // [ads_start[m], ads_start[m+1]); the adsorbate structure factor
  // (StoreTotalChargeAdsorbates) and net charge follow them
  int ads_nmol = 0, ads_n = 0;
  std::vector<int32_t> ads_start, ads_type;
  std::vector<double> ads_xyz;
  double net_q_ads = 0.0;
  int32_t *d_ads_mol = nullptr;
  size_t ads_mol_cap = 0;
  double2 *d_sads = nullptr;
  // tabulated grids in use (FrameworkModel GRID): per pseudo-atom VDW grids
  // and the Coulomb grid, borrowed from their owners
  std::vector<rhg_grid *> vdw_grids;
  rhg_grid *coulomb_grid = nullptr;
  // blocked pockets of the Widom path (device copy of the centres)
  PocketArgs pockets = {};
  double *d_pockets = nullptr;
  std::map<std::string, std::pair<int64_t, double>> stats;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
};

struct rhg_grid {
  rhg_ctx *ctx = nullptr;
  int device = 0;                // freeing needs no live context
  int n[3];
  double size[3], shift[3], delta[3];
  double ubox[9], uinv[9], ulen[3];
  int bc;
  int fp64;
  int coulomb = 0;               // Ewald-Coulomb grid of a unit charge
  double spacing = 0.0;          // SpacingVDWGrid / SpacingCoulombGrid
  double ewald_precision = -1.0; // of a Coulomb grid
  int n_cells[3] = {1, 1, 1};    // NumberOfUnitCells of the system
  double exec_pairs = 0.0;       // (node, atom or image) evaluations of the build
  void *d = nullptr;
  size_t bytes = 0;
};

namespace {

InterpArgs interp_args(const rhg_grid *g, double charge) {
  InterpArgs a;
  memset(&a, 0, sizeof(a));
  for (int d = 0; d < 3; ++d) {
    a.n[d] = g->n[d];
    a.size[d] = g->size[d];
    a.shift[d] = g->shift[d];
    a.ulen[d] = g->ulen[d];
  }
  memcpy(a.ubox, g->ubox, sizeof(a.ubox));
  memcpy(a.uinv, g->uinv, sizeof(a.uinv));
  a.bc = g->bc;
  a.fp64 = g->fp64;
  a.coulomb = g->coulomb;
  a.charge = charge;
  a.grid = g->d;
  return a;
}

rhg_grid *vdw_grid_of(const rhg_ctx *c, int type) {
  return type >= 0 && type < (int)c->vdw_grids.size() ? c->vdw_grids[type] : nullptr;
}

int fail(rhg_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPC(ctx, call)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((ctx), RHG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr size_t kArenaBlock = (size_t)2 << 20;

// a table of the context, in its arena (256-byte aligned)
int dalloc(rhg_ctx *c, void **p, size_t bytes) {
  *p = nullptr;
  bytes = (std::max<size_t>(bytes, 8) + 255) & ~(size_t)255;
  Arena &A = c->arena;
  if (bytes >= kArenaBlock / 2) {
    hipError_t e = dev_malloc(A.pool, p, bytes);
    if (e != hipSuccess) return fail(c, RHG_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    A.large.push_back({(char *)*p, bytes});
    c->footprint += bytes;
    return RHG_OK;
  }
  if (A.small.empty() || A.used + bytes > A.small.back().second) {
    void *blk = nullptr;
    hipError_t e = dev_malloc(A.pool, &blk, kArenaBlock);
    if (e != hipSuccess) return fail(c, RHG_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    A.small.push_back({(char *)blk, kArenaBlock});
    A.used = 0;
    c->footprint += kArenaBlock;
  }
  *p = A.small.back().first + A.used;
  A.used += bytes;
  return RHG_OK;
}

// releases a table of the context: large buffers at once, carved ones with
// the context
void dfree(rhg_ctx *c, void *p) {
  if (!p) return;
  Arena &A = c->arena;
  for (size_t i = 0; i < A.large.size(); ++i)
    if (A.large[i].first == (char *)p) {
      dev_free(A.pool, p, A.large[i].second);
      c->footprint -= A.large[i].second;
      A.large.erase(A.large.begin() + (long)i);
      return;
    }
}

void arena_release(rhg_ctx *c) {
  rhg_pool *P = c->arena.pool;
  for (auto &b : c->arena.small) dev_free(P, b.first, b.second);
  for (auto &b : c->arena.large) dev_free(P, b.first, b.second);
  c->arena = Arena();
  c->arena.pool = P;
}

void fill(double *p, int64_t n, double v, hipStream_t st);

// grow-only named work buffer, in the attached workspace if there is one
// (the context's own counters always stay with the context).  zeroed: a
// (re)allocated buffer is cleared in full, headroom included, on the
// context's stream (buffers whose users keep them cleared, such as the
// per-cell counts of the binning, rely on every byte starting at zero)
void *work(rhg_ctx *c, const char *name, size_t bytes, bool zeroed = false) {
  const bool shared = c->ws && strcmp(name, "pair_counter") != 0;
  DevBuf &b = shared ? c->ws->bufs[name] : c->bufs[name];
  if (b.bytes >= bytes && b.p) return b.p;
  size_t want = std::max<size_t>(bytes, 256);
  if (b.p) {
    // regrowth: leave headroom so that a slowly growing size settles
    want = std::max(want, b.bytes + b.bytes / 2);
    b.retired.push_back({b.p, b.bytes});
  }
  want = (want + 255) & ~(size_t)255;
  b.p = nullptr;
  b.bytes = 0;
  if (shared)
    c->ws->gen++;
  else
    c->gen++;
  // the context's own buffers come from its pool, a workspace's from hipMalloc
  if (dev_malloc(shared ? nullptr : c->arena.pool, &b.p, want) != hipSuccess) {
    b.p = nullptr;
    return nullptr;
  }
  b.bytes = want;
  if (zeroed) fill((double *)b.p, (int64_t)(want / sizeof(double)), 0.0, c->stream); // all-zero bits
  return b.p;
}

hipEvent_t get_event(rhg_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hipEventCreate(&e);
  return e;
}

bool timing_on(const rhg_ctx *c, const char *nm) {
  return c->prof == 1 || (c->prof == 2 && std::strcmp(nm, "site_energy") == 0);
}

// brackets the launches of one named kernel with events when profiling;
// with a list tag the time also counts under "<name>:<tag>"
struct Timed {
  rhg_ctx *c;
  const char *name;
  const char *tag;
  hipEvent_t a = nullptr;
  bool on;
  Timed(rhg_ctx *ctx, const char *nm, const char *tg = nullptr)
      : c(ctx), name(nm), tag(tg), on(timing_on(ctx, nm)) {
    if (on) {
      a = get_event(c);
      hipEventRecord(a, c->stream);
    }
  }
  ~Timed() {
    if (on) {
      hipEvent_t b = get_event(c);
      hipEventRecord(b, c->stream);
      c->pending.push_back({name, tag ? std::string(name) + ":" + tag : std::string(), a, b});
    }
  }
};

void drain_profile(rhg_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0.0f;
    hipEventSynchronize(p.b);
    hipEventElapsedTime(&ms, p.a, p.b);
    auto &s = c->stats[p.name];
    s.first += 1;
    s.second += ms;
    if (!p.stage.empty()) {
      auto &t = c->stats[p.stage];
      t.first += 1;
      t.second += ms;
    }
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
}

bool considered_charged(double q) { return fabs(q) > 1e-10; }

// LENNARD_JONES PotentialValue at r^2 (src/potentials.c:5867-5877)
double lj_ref(double arg1, double sigma, double arg3, double rr) {
  double arg2 = sigma * sigma;
  double t = arg2 / rr;
  double rri3 = (t * t) * t;
  return 4.0 * arg1 * (rri3 * (rri3 - 1.0)) - arg3;
}

// InitializeEwald (automatic), SetupKVectors, EwaldEnergyIon
// (src/ewald.c:200-260, 532-600, 615-742)
void ewald_setup(rhg_ctx *c, double precision) {
  c->ewald_precision = precision;
  double eps = std::min(fabs(precision), 0.5);
  double rc = c->rc_cc;
  double tol = sqrt(fabs(log(eps * rc)));
  c->alpha = sqrt(fabs(log(eps * rc * tol))) / rc;
  double tol1 = sqrt(-log(eps * rc * ((2.0 * tol * c->alpha) * (2.0 * tol * c->alpha))));
  c->kx = nint((double)0.25 + c->props[0] * c->alpha * tol1 / M_PI);
  c->ky = nint((double)0.25 + c->props[1] * c->alpha * tol1 / M_PI);
  c->kz = nint((double)0.25 + c->props[2] * c->alpha * tol1 / M_PI);
  int m = std::max(c->kx, std::max(c->ky, c->kz));
  c->recip_cutoff = (1.05 * m) * (1.05 * m);
  const double *inv = c->cell.inv;
  double volume = c->props[6];
  double alpha = c->alpha;
  c->kfactor.clear();
  c->kidx.clear();
  int ny = 2 * c->ky + 1;
  c->row_off.assign((size_t)(c->kx + 1) * ny, 0);
  c->row_kz.assign((size_t)(c->kx + 1) * ny, -1);
  int nvec = 0;
  for (int ii = 0; ii <= c->kx; ii++) {
    double rk2x = 2.0 * M_PI * ii * inv[AX];
    double rk2y = 2.0 * M_PI * ii * inv[BX];
    double rk2z = 2.0 * M_PI * ii * inv[CX];
    for (int jj = -c->ky; jj <= c->ky; jj++) {
      double rk1x = rk2x + 2.0 * M_PI * jj * inv[AY];
      double rk1y = rk2y + 2.0 * M_PI * jj * inv[BY];
      double rk1z = rk2z + 2.0 * M_PI * jj * inv[CY];
      int row = ii * ny + (jj + c->ky);
      c->row_off[row] = nvec;
      int kzm = -1;
      for (int kk = -c->kz; kk <= c->kz; kk++) {
        double rkx = rk1x + 2.0 * M_PI * kk * inv[AZ];
        double rky = rk1y + 2.0 * M_PI * kk * inv[BZ];
        double rkz = rk1z + 2.0 * M_PI * kk * inv[CZ];
        double ksqr = ii * ii + jj * jj + kk * kk;
        if ((ksqr != 0) && (ksqr < c->recip_cutoff)) {
          double rksqr = rkx * rkx + rky * rky + rkz * rkz;
          double f;
          if (ii == 0)
            f = c->u.C * (2.0 * M_PI / volume) * exp((-0.25 / (alpha * alpha)) * rksqr) / rksqr;
          else
            f = c->u.C * (4.0 * M_PI / volume) * exp((-0.25 / (alpha * alpha)) * rksqr) / rksqr;
          c->kfactor.push_back(f);
          c->kidx.push_back(ii);
          c->kidx.push_back(jj);
          c->kidx.push_back(kk);
          kzm = std::max(kzm, std::abs(kk));
          nvec++;
        }
      }
      c->row_kz[row] = kzm;
    }
  }
  c->nk = nvec;
  double energy = 0;
  for (int k = 0; k < nvec; k++) energy += c->kfactor[k] * ((1.0 * 1.0 + 0.0 * 0.0));
  c->uion = -(energy - c->u.C * c->alpha / sqrt(M_PI));
}

int upload(rhg_ctx *c, void *dst, const void *src, size_t bytes) {
  HIPC(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return RHG_OK;
}

// One copy of a staged host buffer, which the context keeps until its stream
// has next been synchronised.
int upload_kept(rhg_ctx *c, void *dst, std::vector<char> &&host) {
  if (host.empty()) return RHG_OK;
  HIPC(c, hipMemcpyAsync(dst, host.data(), host.size(), hipMemcpyHostToDevice, c->stream));
  c->kept.push_back(std::move(host));
  return RHG_OK;
}

int sync_stream(rhg_ctx *c) {
  HIPC(c, hipStreamSynchronize(c->stream));
  c->kept.clear();
  return RHG_OK;
}

// Packs host arrays into one staging buffer, each part 256-byte aligned:
// add() returns the part's offset.
struct Staging {
  std::vector<char> h;
  size_t add(const void *src, size_t bytes) {
    const size_t o = h.size();
    h.resize(o + ((bytes + 255) & ~(size_t)255), 0);
    if (bytes) memcpy(h.data() + o, src, bytes);
    return o;
  }
};

int setup_pseudo_atom(rhg_ctx *c, PseudoAtom &t) {
  std::vector<AtomRec64> h64((size_t)std::max(c->n, 1));
  std::vector<AtomRecMx> hmx((size_t)std::max(c->n, 1));
  std::vector<double> h24((size_t)std::max(c->n, 1)), h96(h24.size()), h384(h24.size());
  std::vector<float4> hpar(h24.size());
  double ei = t.eps_k * c->u.K2E; // src/potentials.c:268-274
  for (int i = 0; i < c->n; i++) {
    double ej = c->eps[i] * c->u.K2E;
    double e = sqrt(ei * ej);              // src/potentials.c:3164-3171
    double sg = 0.5 * (t.sigma + c->sig[i]);
    double shift = c->shifted ? lj_ref(e, sg, 0.0, c->rc2_vdw) : 0.0; // src/potentials.c:5263-5285
    AtomRec64 &r = h64[i];
    r.x = c->x[i];
    r.y = c->y[i];
    r.z = c->z[i];
    r.c4 = 4.0 * e;
    r.s2 = sg * sg;
    r.shift = shift;
    r.q = c->q[i];
    r.pad = 0.0;
    AtomRecMx &m = hmx[i];
    m.x = r.x;
    m.y = r.y;
    m.z = r.z;
    m.c4 = (float)r.c4;
    m.s2 = (float)r.s2;
    m.shift = (float)r.shift;
    m.q = (float)r.q;
    h24[i] = 24.0 * e;
    h96[i] = 96.0 * e;
    h384[i] = 384.0 * e;
    const double s6 = r.s2 * r.s2 * r.s2;
    hpar[i] = make_float4((float)(r.c4 * s6 * s6), (float)(r.c4 * s6), (float)r.shift, (float)r.q);
  }
  // the six record arrays of this type in one table and one copy
  const size_t n = h64.size();
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o64 = 0, omx = o64 + up(n * sizeof(AtomRec64)), o24 = omx + up(n * sizeof(AtomRecMx)),
               o96 = o24 + up(n * sizeof(double)), o384 = o96 + up(n * sizeof(double)),
               opar = o384 + up(n * sizeof(double)), total = opar + up(n * sizeof(float4));
  std::vector<char> host(total, 0);
  memcpy(host.data() + o64, h64.data(), n * sizeof(AtomRec64));
  memcpy(host.data() + omx, hmx.data(), n * sizeof(AtomRecMx));
  memcpy(host.data() + o24, h24.data(), n * sizeof(double));
  memcpy(host.data() + o96, h96.data(), n * sizeof(double));
  memcpy(host.data() + o384, h384.data(), n * sizeof(double));
  memcpy(host.data() + opar, hpar.data(), n * sizeof(float4));
  char *d = nullptr;
  if (int rc = dalloc(c, (void **)&d, total)) return rc;
  t.d64 = (AtomRec64 *)(d + o64);
  t.dmx = (AtomRecMx *)(d + omx);
  t.c24 = (double *)(d + o24);
  t.c96 = (double *)(d + o96);
  t.c384 = (double *)(d + o384);
  t.par = (float4 *)(d + opar);
  return upload_kept(c, d, std::move(host));
}

EnergyArgs energy_args(rhg_ctx *c, int type, double scaling, int do_vdw, int do_cc) {
  EnergyArgs a;
  a.cell = c->cell;
  a.rc2_vdw = c->rc2_vdw;
  a.rc2_cc = c->rc2_cc;
  a.alpha = c->alpha;
  const PseudoAtom &t = c->types[type];
  double chargeA = scaling * t.q;
  a.cqa = c->u.C * chargeA;
  a.n_atoms = c->n;
  a.a64 = t.d64;
  a.amx = t.dmx;
  a.par = t.par;
  a.precision = c->precision;
  a.do_vdw = do_vdw;
  a.do_cc = (do_cc && c->charge_method == RHG_CHARGE_EWALD && considered_charged(t.q) && c->n > 0) ? 1 : 0;
  return a;
}

// The adsorbate atoms as seen by pseudo-atom `type`: positions, LB-mixed LJ
// of the type pair (as for framework atoms) and the atom's charge.
int ads_records(rhg_ctx *c, PseudoAtom &t) {
  if (c->ads_n == 0) return RHG_OK;
  std::vector<AtomRec64> h((size_t)c->ads_n);
  const double ei = t.eps_k * c->u.K2E;
  for (int i = 0; i < c->ads_n; ++i) {
    const PseudoAtom &b = c->types[c->ads_type[i]];
    const double e = sqrt(ei * (b.eps_k * c->u.K2E));
    const double sg = 0.5 * (t.sigma + b.sigma);
    AtomRec64 &r = h[i];
    r.x = c->ads_xyz[3 * i];
    r.y = c->ads_xyz[3 * i + 1];
    r.z = c->ads_xyz[3 * i + 2];
    r.c4 = 4.0 * e;
    r.s2 = sg * sg;
    r.shift = c->shifted ? lj_ref(e, sg, 0.0, c->rc2_vdw) : 0.0;
    r.q = 1.0 * b.q;
    r.pad = 0.0;
  }
  if (t.dads_cap < h.size()) {
    dfree(c, t.dads);
    t.dads = nullptr;
    t.dads_cap = 0;
    if (int rc = dalloc(c, (void **)&t.dads, h.size() * sizeof(AtomRec64))) return rc;
    t.dads_cap = h.size();
  }
  if (int rc = upload(c, t.dads, h.data(), h.size() * sizeof(AtomRec64))) return rc;
  if (int rs_ = sync_stream(c)) return rs_; // h is released on return
  return RHG_OK;
}

AdsArgs ads_args(rhg_ctx *c, int type, int exclude, int do_vdw, int do_cc) {
  AdsArgs a;
  memset(&a, 0, sizeof(a));
  const PseudoAtom &t = c->types[type];
  a.cell = c->cell;
  a.rec = t.dads;
  a.mol = c->d_ads_mol;
  a.n = c->ads_n;
  a.exclude = exclude;
  a.rc2v = c->rc2_vdw;
  a.rc2c = c->rc2_cc;
  a.alpha = c->alpha;
  a.cqa = c->u.C * (1.0 * t.q); // chargeA = CFChargeScaling * Charge1 (src/cbmc.c:687)
  a.do_vdw = do_vdw;
  a.do_cc = (do_cc && c->charge_method == RHG_CHARGE_EWALD && considered_charged(t.q)) ? 1 : 0;
  return a;
}

__global__ void k_fill(double *p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void fill(double *p, int64_t n, double v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

// The k-vectors (indices into RASPA2's k order) in the order k_fourier_part
// visits them: per kx slice the k_y = 0 row (k_z = 0 unless kx = 0, then
// +k_z and -k_z per step), then per k_y > 0 the rows +k_y and -k_y together
// (k_z = 0 of each, then (+k_y, +k_z), (+k_y, -k_z), (-k_y, +k_z),
// (-k_y, -k_z) per step); off[kx] is the first of slice kx, off[kx + 1] the
// count.
void fourier_order(const rhg_ctx *c, std::vector<int32_t> &order, std::vector<int32_t> &off) {
  const int ny = 2 * c->ky + 1;
  order.clear();
  off.assign((size_t)c->kx + 2, 0);
  for (int ii = 0; ii <= c->kx; ++ii) {
    off[ii] = (int32_t)order.size();
    for (int jj = 0; jj <= c->ky; ++jj) {
      if (jj == 0) {
        const int row = ii * ny + c->ky, kzm = c->row_kz[row];
        if (kzm < 0) continue;
        const int o = c->row_off[row];
        if (ii != 0) order.push_back(o + kzm);
        const int ip = o + kzm - (ii == 0 ? 1 : 0), im = o + kzm;
        for (int kk = 1; kk <= kzm; ++kk) {
          order.push_back(ip + kk);
          order.push_back(im - kk);
        }
        continue;
      }
      const int rp = ii * ny + (jj + c->ky), rm = ii * ny + (c->ky - jj);
      const int kzm = c->row_kz[rp];
      if (kzm < 0) continue;
      const int op = c->row_off[rp] + kzm, om = c->row_off[rm] + kzm;
      order.push_back(op);
      order.push_back(om);
      for (int kk = 1; kk <= kzm; ++kk) {
        order.push_back(op + kk);
        order.push_back(op - kk);
        order.push_back(om + kk);
        order.push_back(om - kk);
      }
    }
  }
  off[(size_t)c->kx + 1] = (int32_t)order.size();
}

EwaldDev ewald_dev(rhg_ctx *c) {
  EwaldDev e;
  e.kx = c->kx;
  e.ky = c->ky;
  e.kz = c->kz;
  e.nk = c->nk;
  e.alpha = c->alpha;
  e.uion = c->uion;
  e.net_q_fw = c->net_q_fw;
  e.beta_unused = 0.0;
  e.kfactor = c->d_kfactor;
  e.sf = c->d_sf;
  e.row_off = c->d_row_off;
  e.row_kz = c->d_row_kz;
  e.kst = c->d_kst;
  for (int i = 0; i < kMaxKx + 2; ++i)
    e.kst_off[i] = i < (int)c->kst_off.size() ? c->kst_off[i] : 0;
  e.cell = c->cell;
  e.C = c->u.C;
  return e;
}

int check_type(rhg_ctx *c, int32_t type) {
  if (type < 0 || type >= (int)c->types.size()) return fail(c, RHG_ERR_INVALID, "unknown pseudo-atom type");
  return RHG_OK;
}

// Cell grid and neighbour lists for culled site energies (built lazily).
bool image_lists_apply(const rhg_ctx *c);
double image_box_reach(const rhg_ctx *c);
int image_lists_ready(rhg_ctx *c);
int cell_shift_room(rhg_ctx *c, PseudoAtom &t);

int ensure_cell_grid(rhg_ctx *c) {
  if (c->grid_ready) return RHG_OK;
  hipStream_t st = c->stream;
  int nc[3];
  for (int d = 0; d < 3; ++d) {
    nc[d] = (int)floor(c->props[3 + d] / c->cell_size);
    nc[d] = std::max(1, std::min(nc[d], 64));
  }
  int ncell = nc[0] * nc[1] * nc[2];
  // half the longest diagonal of one cell
  const double *B = c->cell.box;
  double u[3] = {B[AX] / nc[0], B[AY] / nc[0], B[AZ] / nc[0]};
  double v[3] = {B[BX] / nc[1], B[BY] / nc[1], B[BZ] / nc[1]};
  double w[3] = {B[CX] / nc[2], B[CY] / nc[2], B[CZ] / nc[2]};
  double R = 0.0;
  for (int sv = -1; sv <= 1; sv += 2)
    for (int sw = -1; sw <= 1; sw += 2) {
      double d2 = 0.0;
      for (int k = 0; k < 3; ++k) {
        double e = u[k] + sv * v[k] + sw * w[k];
        d2 += e * e;
      }
      R = std::max(R, 0.5 * sqrt(d2));
    }
  double rc = std::max(c->rc_vdw, c->rc_cc);
  double reach = rc + R + 1e-6;
  int n = std::max(c->n, 1);
  int rc_;
  if (!c->d_ax) {
    // the atom positions in one block and one copy
    Staging xyz;
    const size_t ox = xyz.add(c->x.data(), c->n * sizeof(double)), oy = xyz.add(c->y.data(), c->n * sizeof(double)),
                 oz = xyz.add(c->z.data(), c->n * sizeof(double));
    char *d = nullptr;
    if ((rc_ = dalloc(c, (void **)&d, std::max<size_t>(xyz.h.size(), 3 * 256)))) return rc_;
    c->d_ax = (double *)(d + ox);
    c->d_ay = (double *)(d + (c->n ? oy : 256));
    c->d_az = (double *)(d + (c->n ? oz : 512));
    if ((rc_ = upload_kept(c, d, std::move(xyz.h)))) return rc_;
  }
  c->cg.nc[0] = nc[0];
  c->cg.nc[1] = nc[1];
  c->cg.nc[2] = nc[2];
  c->cg.ncell = ncell;
  c->cg.img_off = nullptr;
  c->cg.img = nullptr;
  c->img_ready = false;
  c->cell_reach = reach;
  c->cell_halfdiag = R;
  int32_t *counts = (int32_t *)work(c, "nb_counts", (size_t)ncell * sizeof(int32_t));
  dfree(c, c->d_nb_off);
  c->d_nb_off = nullptr;
  if ((rc_ = dalloc(c, (void **)&c->d_nb_off, (size_t)(ncell + 1) * sizeof(int32_t)))) return rc_;
  if (!counts) return fail(c, RHG_ERR_OOM, "cell counts");
  // mixed precision with image lists: both lists in one pass pair
  if (c->precision != RHG_PREC_FP64 && c->use_images && image_lists_apply(c)) {
    int32_t *icounts = (int32_t *)work(c, "img_counts", (size_t)ncell * sizeof(int32_t));
    int32_t *tot = (int32_t *)work(c, "list_totals", 2 * sizeof(int32_t));
    if (!icounts || !tot) return fail(c, RHG_ERR_OOM, "cell counts");
    dfree(c, c->d_img_off);
    c->d_img_off = nullptr;
    if ((rc_ = dalloc(c, (void **)&c->d_img_off, (size_t)(ncell + 1) * sizeof(int32_t)))) return rc_;
    const double box_reach = image_box_reach(c);
    launch_cell_lists(c->cell, c->cg, reach, c->props + 3, box_reach, c->d_ax, c->d_ay, c->d_az, c->n, counts,
                      icounts, nullptr, nullptr, nullptr, nullptr, 0, st);
    launch_scan2(counts, c->d_nb_off, icounts, c->d_img_off, ncell, tot, st);
    int32_t totals[2] = {0, 0};
    HIPC(c, hipMemcpyAsync(totals, tot, sizeof(totals), hipMemcpyDeviceToHost, st));
    if (int rs = sync_stream(c)) return rs;
    dfree(c, c->d_nb_idx);
    c->d_nb_idx = nullptr;
    dfree(c, c->d_img);
    c->d_img = nullptr;
    if ((rc_ = dalloc(c, (void **)&c->d_nb_idx, (size_t)std::max(totals[0], 1) * sizeof(int32_t)))) return rc_;
    if ((rc_ = dalloc(c, (void **)&c->d_img, (size_t)std::max(totals[1], 1) * sizeof(float4)))) return rc_;
    launch_cell_lists(c->cell, c->cg, reach, c->props + 3, box_reach, c->d_ax, c->d_ay, c->d_az, c->n, counts,
                      icounts, c->d_nb_off, c->d_nb_idx, c->d_img_off, c->d_img, 1, st);
    HIPC(c, hipGetLastError());
    c->cg.nb_off = c->d_nb_off;
    c->cg.nb_idx = c->d_nb_idx;
    c->grid_ready = true;
    c->gen++;
    return image_lists_ready(c);
  }
  launch_cell_neighbours(c->cell, nc[0], nc[1], nc[2], reach, c->d_ax, c->d_ay, c->d_az, c->n,
                         counts, nullptr, nullptr, 0, st);
  launch_scan_counts(counts, ncell, c->d_nb_off, st);
  int32_t total = 0;
  HIPC(c, hipMemcpyAsync(&total, c->d_nb_off + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (int rs = sync_stream(c)) return rs;
  dfree(c, c->d_nb_idx);
  c->d_nb_idx = nullptr;
  if ((rc_ = dalloc(c, (void **)&c->d_nb_idx, (size_t)std::max(total, 1) * sizeof(int32_t)))) return rc_;
  launch_cell_neighbours(c->cell, nc[0], nc[1], nc[2], reach, c->d_ax, c->d_ay, c->d_az, c->n,
                         counts, c->d_nb_off, c->d_nb_idx, 1, st);
  HIPC(c, hipGetLastError());
  c->cg.nb_off = c->d_nb_off;
  c->cg.nb_idx = c->d_nb_idx;
  c->grid_ready = true;
  c->gen++;
  return RHG_OK;
}

// Image lists of the cell grid for the mixed-precision kernels.  They hold
// every image within reach, which equals the minimum-image convention when
// every perpendicular width is at least twice the cut-off; the FP32 offsets
// stay below 31 A (rounding below 1e-6 A), and the 27 lattice shifts must
// cover the reach.  Otherwise the lists stay off and the mixed kernels take the
// minimum image per pair.
bool image_lists_apply(const rhg_ctx *c) {
  const double wmin = std::min(c->props[3], std::min(c->props[4], c->props[5]));
  const double rc = std::max(c->rc_vdw, c->rc_cc);
  return wmin >= 2.0 * rc && c->cell_reach + c->cell_halfdiag < 31.0 && c->cell_reach < 1.5 * wmin;
}

// rectangular cells: the images within the cut-off of the cell's box (a
// 1e-3 A margin covers the FP32 rounding of the offsets, below 1e-5 A)
double image_box_reach(const rhg_ctx *c) { return std::max(c->rc_vdw, c->rc_cc) + 1e-3; }

// The image lists of d_img_off / d_img are complete: publish them and make
// room for the per-cell sums that depend on them.
int image_lists_ready(rhg_ctx *c) {
  c->img_ready = true;
  c->cg.img_off = c->d_img_off;
  c->cg.img = c->d_img;
  c->img_version++;
  c->gen++;
  for (PseudoAtom &t : c->types)
    if (int rc = cell_shift_room(c, t)) return rc;
  return RHG_OK;
}

int ensure_image_list(rhg_ctx *c) {
  if (c->img_ready) return RHG_OK;
  c->img_ready = true;
  if (!image_lists_apply(c)) return RHG_OK;
  hipStream_t st = c->stream;
  const int ncell = c->cg.ncell;
  int32_t *counts = (int32_t *)work(c, "img_counts", (size_t)ncell * sizeof(int32_t));
  if (!counts) return fail(c, RHG_ERR_OOM, "image counts");
  dfree(c, c->d_img_off);
  c->d_img_off = nullptr;
  int rc_;
  if ((rc_ = dalloc(c, (void **)&c->d_img_off, (size_t)(ncell + 1) * sizeof(int32_t)))) return rc_;
  const double box_reach = image_box_reach(c);
  launch_image_list(c->cell, c->cg, c->cell_reach, c->props + 3, box_reach, c->d_ax, c->d_ay, c->d_az,
                    counts, nullptr, nullptr, 0, st);
  launch_scan_counts(counts, ncell, c->d_img_off, st);
  int32_t total = 0;
  HIPC(c, hipMemcpyAsync(&total, c->d_img_off + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (int rs = sync_stream(c)) return rs;
  dfree(c, c->d_img);
  c->d_img = nullptr;
  if ((rc_ = dalloc(c, (void **)&c->d_img, (size_t)std::max(total, 1) * sizeof(float4)))) return rc_;
  launch_image_list(c->cell, c->cg, c->cell_reach, c->props + 3, box_reach, c->d_ax, c->d_ay, c->d_az,
                    counts, c->d_img_off, c->d_img, 1, st);
  HIPC(c, hipGetLastError());
  return image_lists_ready(c);
}

// Room for the per-cell shift sums of pseudo-atom t over the current image
// lists (allocated with the lists and with each new pseudo-atom, so that
// the sums can be formed later without allocating).
int cell_shift_room(rhg_ctx *c, PseudoAtom &t) {
  if (!c->cg.img || t.cell_shift_cap >= c->cg.ncell) return RHG_OK;
  dfree(c, t.cell_shift);
  t.cell_shift = nullptr;
  t.cell_shift_cap = 0;
  if (int rc = dalloc(c, (void **)&t.cell_shift, (size_t)c->cg.ncell * sizeof(double))) return rc;
  t.cell_shift_cap = c->cg.ncell;
  t.cell_shift_version = 0;
  return RHG_OK;
}

// The per-cell shift sums over the current image lists of the pseudo-atom
// whose FP64 records are a64 (*out = null when a64 is no pseudo-atom's),
// formed on first use and after the lists are rebuilt; may_alloc = false
// never allocates (the room exists unless an allocation failed).
int cell_shifts(rhg_ctx *c, const AtomRec64 *a64, bool may_alloc, const double **out) {
  *out = nullptr;
  for (PseudoAtom &t : c->types) {
    if (t.d64 != a64) continue;
    if (may_alloc)
      if (int rc = cell_shift_room(c, t)) return rc;
    if (t.cell_shift_cap < c->cg.ncell) return RHG_OK;
    if (t.cell_shift_version != c->img_version) {
      launch_img_shift_sums(c->cg, t.d64, t.cell_shift, c->stream);
      HIPC(c, hipGetLastError());
      t.cell_shift_version = c->img_version;
    }
    *out = t.cell_shift;
    return RHG_OK;
  }
  return RHG_OK;
}

// Sites per tile of a list's image-list kernel: the launch-shape knob for
// LJ-only lists; one site per lane where Coulomb terms are summed (their
// pair loop holds more registers per site, and 64-site tiles give sparse
// lists, such as first beads and charged beads, more waves).
int list_tile_sites(const rhg_ctx *c, const EnergyArgs &a) {
  if (a.do_cc && !a.do_vdw) return 32; // Coulomb-only lists: two 32-site tiles per wave
  if (a.do_cc) return 64;
  return c->tile_sites;
}

int site_energy_pairs(rhg_ctx *c, const EnergyArgs &a, const SiteList &s, const EnergyOut &o,
                      const char *tag, const BinWork *given = nullptr);

// Site energies of one site list of pseudo-atom `type`: terms with a grid in
// use are interpolated (src/framework_energy.c:73-76, 197-200: the Coulomb
// grid value times the pseudo-atom's charge, unscaled), the others summed
// over the framework atoms.
int site_energy(rhg_ctx *c, int type, const EnergyArgs &a, const SiteList &s, const EnergyOut &o,
                const char *tag, const BinWork *given = nullptr) {
  if (s.n <= 0) return RHG_OK;
  rhg_grid *gv = a.do_vdw ? vdw_grid_of(c, type) : nullptr;
  rhg_grid *gc = a.do_cc ? c->coulomb_grid : nullptr;
  if (!gv && !gc) return site_energy_pairs(c, a, s, o, tag, given);
  EnergyArgs rest = a;
  rest.do_vdw = a.do_vdw && !gv;
  rest.do_cc = a.do_cc && !gc;
  const bool merge = rest.do_vdw || rest.do_cc;
  if (merge)
    if (int rc = site_energy_pairs(c, rest, s, o, tag, given)) return rc;
  InterpArgs iv, ic;
  if (gv) iv = interp_args(gv, 1.0);
  if (gc) ic = interp_args(gc, c->types[type].q);
  Timed tm(c, "grid_interp");
  launch_site_grid(gv ? &iv : nullptr, gc ? &ic : nullptr, s, o, merge ? 1 : 0, c->stream);
  return RHG_OK;
}

// The culled site-energy kernel of one list, timed (when profiling) by
// events its own dispatch records: the kernel alone, under "site_energy"
// and "site_energy:<tag>".
void site_energy_launch(rhg_ctx *c, const EnergyArgs &a, const CellGrid &g, const SiteList &s,
                        const BinWork &b, const EnergyOut &o, const char *tag) {
  LaunchEvents ev;
  const bool timed = timing_on(c, "site_energy");
  if (timed) {
    ev.start = get_event(c);
    ev.stop = get_event(c);
  }
  launch_site_energy_culled(a, g, s, b, o, c->stream, ev);
  if (timed) c->pending.push_back({"site_energy", std::string("site_energy:") + tag, ev.start, ev.stop});
}

// Site energies of one site list, culled through the cell grid when enabled.
// `tag` keeps the binning scratch of concurrent lists apart; `given`
// (nullable) is binning scratch the caller provides, for lists whose cell
// grid (and image lists) already exist.
int site_energy_pairs(rhg_ctx *c, const EnergyArgs &a, const SiteList &s, const EnergyOut &o,
                      const char *tag, const BinWork *given) {
  if (s.n <= 0) return RHG_OK;
  if (!c->cull || c->n == 0) {
    Timed t(c, "site_energy", tag);
    launch_site_energy(a, s, o, c->stream);
    return RHG_OK;
  }
  if (int rc = ensure_cell_grid(c)) return rc;
  if (c->precision != RHG_PREC_FP64 && c->use_images)
    if (int rc = ensure_image_list(c)) return rc;
  CellGrid g = c->cg;
  if (!c->use_images) {
    g.img_off = nullptr;
    g.img = nullptr;
  }
  // shifted LJ on the image lists: the clamped form with the cells' shift sums
  EnergyArgs ae = a;
  ae.cell_shift = nullptr;
  if (a.do_vdw && c->shifted && a.precision != RHG_PREC_FP64 && g.img_off && a.par)
    if (int rc = cell_shifts(c, a.a64, given == nullptr, &ae.cell_shift)) return rc;
  if (given) {
    BinWork b = *given;
    b.tile_sites = list_tile_sites(c, a);
    b.waves = c->waves;
    if (!(a.precision != RHG_PREC_FP64 && g.img_off && a.par)) b.fix = nullptr;
    {
      Timed tm(c, "bin");
      launch_bin_sites(c->cell, c->cg, s, b, c->stream);
    }
    site_energy_launch(c, ae, g, s, b, o, tag);
    return RHG_OK;
  }
  const int ncell = c->cg.ncell;
  std::string t = tag;
  BinWork b;
  b.tile_sites = list_tile_sites(c, a);
  b.waves = c->waves;
  b.cell_of = (int32_t *)work(c, (t + "_cellof").c_str(), (size_t)s.n * sizeof(int32_t));
  b.perm = (int32_t *)work(c, (t + "_perm").c_str(), (size_t)s.n * sizeof(int32_t));
  // cleared in full when allocated; each binning's scan clears what it used
  b.count = (int32_t *)work(c, (t + "_count").c_str(), (size_t)ncell * sizeof(int32_t), true);
  b.counts_zero = 1;
  b.rank = (int32_t *)work(c, (t + "_rank").c_str(), (size_t)s.n * sizeof(int32_t));
  b.start = (int32_t *)work(c, (t + "_start").c_str(), (size_t)(ncell + 1) * sizeof(int32_t));
  b.tile_off = (int32_t *)work(c, (t + "_tiles").c_str(), (size_t)(ncell + 1) * sizeof(int32_t));
  // the image-list kernels read the sites' sorted FP32 offsets
  b.fix = nullptr;
  if (a.precision != RHG_PREC_FP64 && g.img_off && a.par) {
    b.fix = (float4 *)work(c, (t + "_fix").c_str(), (size_t)s.n * sizeof(float4));
    b.fixu = (float4 *)work(c, (t + "_fixu").c_str(), (size_t)s.n * sizeof(float4));
    if (!b.fixu) return fail(c, RHG_ERR_OOM, "binning scratch");
    b.redo = (int2 *)work(c, (t + "_redo").c_str(), (size_t)s.n * sizeof(int2));
    b.n_redo = (int32_t *)work(c, (t + "_nredo").c_str(), 256);
    if (!b.fix || !b.redo || !b.n_redo) return fail(c, RHG_ERR_OOM, "binning scratch");
  }
  bool fresh = c->bufs["pair_counter"].p == nullptr;
  b.pairs = (unsigned long long *)work(c, "pair_counter", 3 * kPairLanes * sizeof(unsigned long long));
  if (fresh && b.pairs) fill((double *)b.pairs, 3 * kPairLanes, 0.0, c->stream); // all-zero bits
  if (!b.cell_of || !b.perm || !b.count || !b.rank || !b.start || !b.tile_off || !b.pairs)
    return fail(c, RHG_ERR_OOM, "binning scratch");
  {
    Timed tm(c, "bin");
    launch_bin_sites(c->cell, c->cg, s, b, c->stream);
  }
  site_energy_launch(c, ae, g, s, b, o, tag);
  return RHG_OK;
}

// Structure factor sum_j q_j exp(i k.r_j) of the given charges (fractional
// coordinates times 2 pi) in RASPA2's k order into out (src/ewald.c:1330-1345)
int structure_factor(rhg_ctx *c, const std::vector<double> &fx, const std::vector<double> &fy,
                     const std::vector<double> &fz, const std::vector<double> &qq, double2 *out) {
  const int nc = (int)qq.size();
  const size_t bn = (size_t)nc * sizeof(double);
  Staging in;
  const size_t ox = in.add(fx.data(), bn), oy = in.add(fy.data(), bn), oz = in.add(fz.data(), bn),
               oq = in.add(qq.data(), bn);
  char *d_in = (char *)work(c, "sf_in", std::max<size_t>(in.h.size(), 256));
  const size_t tab_n = (size_t)std::max(nc, 1) * ((c->kx + 1) + (2 * c->ky + 1) + 1);
  double2 *d_tab = (double2 *)work(c, "sf_tab", tab_n * sizeof(double2));
  if (!d_in || !d_tab) return fail(c, RHG_ERR_OOM, "structure factor buffers");
  if (int rc = upload_kept(c, d_in, std::move(in.h))) return rc;
  {
    Timed t(c, "sf");
    launch_structure_factor(ewald_dev(c), (const double *)(d_in + ox), (const double *)(d_in + oy),
                            (const double *)(d_in + oz), (const double *)(d_in + oq), nc, d_tab, out,
                            c->stream);
  }
  HIPC(c, hipGetLastError());
  return RHG_OK;
}

// The framework structure factor of the charged atoms (src/ewald.c:1330-1345)
// and the Fourier kernel's k records, formed once, on the first Fourier use.
int ensure_fourier(rhg_ctx *c) {
  if (c->sf_ready || c->charge_method != RHG_CHARGE_EWALD) return RHG_OK;
  std::vector<double> fx, fy, fz, qq;
  const double *inv = c->cell.inv;
  for (int i = 0; i < c->n; ++i) {
    if (!considered_charged(c->q[i])) continue;
    double p[3] = {c->x[i], c->y[i], c->z[i]};
    double a = inv[AX] * p[0] + inv[BX] * p[1] + inv[CX] * p[2];
    double b = inv[AY] * p[0] + inv[BY] * p[1] + inv[CY] * p[2];
    double d = inv[AZ] * p[0] + inv[BZ] * p[1] + inv[CZ] * p[2];
    fx.push_back(a * (2.0 * M_PI));
    fy.push_back(b * (2.0 * M_PI));
    fz.push_back(d * (2.0 * M_PI));
    qq.push_back(c->q[i]);
  }
  if (int rc = structure_factor(c, fx, fy, fz, qq, c->d_sf)) return rc;
  launch_kstream(c->d_kfactor, c->d_sf, c->d_korder, c->korder_n, c->d_kst, c->stream);
  HIPC(c, hipGetLastError());
  c->sf_ready = true;
  return RHG_OK;
}

void frac_2pi(const rhg_ctx *c, const double p[3], double f[3]) {
  const double *inv = c->cell.inv;
  f[0] = (inv[AX] * p[0] + inv[BX] * p[1] + inv[CX] * p[2]) * (2.0 * M_PI);
  f[1] = (inv[AY] * p[0] + inv[BY] * p[1] + inv[CY] * p[2]) * (2.0 * M_PI);
  f[2] = (inv[AZ] * p[0] + inv[BZ] * p[1] + inv[CZ] * p[2]) * (2.0 * M_PI);
}

} // namespace

// ======================================================================
// Binning scratch of rhg_framework_energy_at_positions_noalloc: a fixed
// device array of the library, so that the entry point allocates nothing.
// Layout (each part 256-byte aligned): site x, y, z (FP64), cell of each
// site, sorted order, sorted FP32 offsets, then per cell the site counts,
// the sites' ranks (in a slot of per-cell size), starts and tile offsets,
// then the list of sites to recompute exactly and its count.  Each GPU holds its own copy; its
// address comes from hipGetSymbolAddress once per device, at context
// creation with that device current.
// ======================================================================
namespace {
constexpr int64_t kNaSites = (int64_t)1 << 18;   // positions per call
constexpr int64_t kNaCells = 64 * 64 * 64;       // the cell grid's largest split
constexpr size_t na_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaXyz = na_round(kNaSites * sizeof(double));
constexpr size_t kNaI32 = na_round(kNaSites * sizeof(int32_t));
constexpr size_t kNaFix = na_round(kNaSites * sizeof(float4));
constexpr size_t kNaCell = na_round((kNaCells + 1) * sizeof(int32_t));
constexpr size_t kNaBytes = 3 * kNaXyz + 4 * kNaI32 + kNaFix + 4 * kNaCell + 256;
static_assert(kNaCell >= kNaI32, "the per-site ranks use a per-cell slot");
constexpr int kNaDevices = 64;
} // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
std::mutex na_mu;
unsigned char *na_ptr[kNaDevices] = {};   // device address of `scratch`, per device
hipError_t na_err[kNaDevices] = {};

// called with `device` current
void noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNaDevices) return;
  std::lock_guard<std::mutex> lk(na_mu);
  if (na_ptr[device] || na_err[device] != hipSuccess) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p)
    na_ptr[device] = (unsigned char *)p;
  else
    na_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
}
} // namespace

// ======================================================================
// C ABI
// ======================================================================
extern "C" {

const char *rhg_version(void) { return "rhg 0.1 (gfx950)"; }

const char *rhg_last_error(const rhg_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

} // extern "C"

namespace {
struct CtxDeleter {
  void operator()(rhg_ctx *c) const { rhg_destroy(c); }
};

// rhg_create; with `stream` given the context borrows it instead of creating
// its own, and with `ws` given its set-up scratch lives in that workspace
int create_impl(const rhg_framework *fw, const rhg_cell *cellp, const rhg_forcefield *ff, int device,
                hipStream_t stream, rhg_workspace *ws, rhg_pool *pool, rhg_ctx **out) {
  if (!fw || !cellp || !ff || !out) return RHG_ERR_INVALID;
  *out = nullptr;
  if (fw->n_atoms < 0 || (fw->n_atoms > 0 && (!fw->x || !fw->y || !fw->z || !fw->eps_k || !fw->sigma)))
    return RHG_ERR_INVALID;
  if (cellp->boundary != RHG_BC_RECTANGULAR && cellp->boundary != RHG_BC_TRICLINIC) return RHG_ERR_INVALID;
  if (ff->cutoff_vdw <= 0.0 || ff->temperature <= 0.0) return RHG_ERR_INVALID;
  if (ff->charge_method != RHG_CHARGE_NONE && ff->charge_method != RHG_CHARGE_EWALD) return RHG_ERR_UNSUPPORTED;
  if ((ws && ws->device != device) || (pool && pool->device != device)) return RHG_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return RHG_ERR_HIP;
  noalloc_scratch_init(device);
  hipStream_t st = stream;
  if (!st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return RHG_ERR_HIP;
  std::unique_ptr<rhg_ctx, CtxDeleter> c(new rhg_ctx());
  c->device = device;
  c->stream = st;
  c->own_stream = stream == nullptr;
  c->arena.pool = pool;
  if (ws) {
    // the workspace's previous user finishes first (set-up runs on this stream)
    if (ws->owner && ws->owner != c.get()) {
      if (hipStreamSynchronize(ws->owner->stream) != hipSuccess) return RHG_ERR_HIP;
      ws->owner->ws = nullptr;
    }
    ws->owner = c.get();
    c->ws = ws;
  }
  c->u = make_units();
  c->cell = make_cell(cellp->box, cellp->boundary);
  cell_properties(c->cell.box, c->props);
  c->n = fw->n_atoms;
  c->x.assign(fw->x, fw->x + c->n);
  c->y.assign(fw->y, fw->y + c->n);
  c->z.assign(fw->z, fw->z + c->n);
  c->eps.assign(fw->eps_k, fw->eps_k + c->n);
  c->sig.assign(fw->sigma, fw->sigma + c->n);
  c->q.assign((size_t)c->n, 0.0);
  if (fw->charge)
    for (int i = 0; i < c->n; ++i) c->q[i] = fw->charge[i];
  c->mass_total = 0.0;
  for (int i = 0; i < c->n; ++i) c->mass_total += fw->mass ? fw->mass[i] : 12.0;
  c->rc_vdw = ff->cutoff_vdw;
  c->rc2_vdw = ff->cutoff_vdw * ff->cutoff_vdw;
  c->shifted = ff->shift_vdw ? 1 : 0;
  c->charge_method = ff->charge_method;
  c->rc_cc = ff->cutoff_charge > 0.0 ? ff->cutoff_charge : ff->cutoff_vdw;
  c->rc2_cc = c->rc_cc * c->rc_cc;
  c->temperature = ff->temperature;
  c->beta = 1.0 / (c->u.KB * c->temperature); // src/input.c:2071
  if (c->charge_method == RHG_CHARGE_EWALD) {
    ewald_setup(c.get(), ff->ewald_precision > 0.0 ? ff->ewald_precision : 1e-6);
    int rc;
    size_t nk = (size_t)std::max(c->nk, 1);
    if (c->kz > 64 || c->kx > kMaxKx) return RHG_ERR_UNSUPPORTED; // k ranges of the row kernel
    // the k-vector records in the Fourier kernel's visiting order
    std::vector<int32_t> order;
    fourier_order(c.get(), order, c->kst_off);
    if ((int)order.size() != c->nk) return fail(c.get(), RHG_ERR_INVALID, "k-vector visiting order");
    order.resize(order.size() + 8, 0); // read-ahead padding
    // the host-built k tables in one block and one copy
    Staging kt;
    const size_t okf = kt.add(c->kfactor.data(), c->kfactor.size() * sizeof(double)),
                 ooff = kt.add(c->row_off.data(), c->row_off.size() * sizeof(int32_t)),
                 okz = kt.add(c->row_kz.data(), c->row_kz.size() * sizeof(int32_t)),
                 oord = kt.add(order.data(), order.size() * sizeof(int32_t));
    char *d_kt = nullptr;
    if ((rc = dalloc(c.get(), (void **)&d_kt, kt.h.size()))) return rc;
    c->d_kfactor = (double *)(d_kt + okf);
    c->d_row_off = (int32_t *)(d_kt + ooff);
    c->d_row_kz = (int32_t *)(d_kt + okz);
    const int32_t *d_order = (const int32_t *)(d_kt + oord);
    if ((rc = dalloc(c.get(), (void **)&c->d_sf, nk * sizeof(double2)))) return rc;
    if ((rc = dalloc(c.get(), (void **)&c->d_kst, order.size() * sizeof(double4)))) return rc;
    if ((rc = upload_kept(c.get(), d_kt, std::move(kt.h)))) return rc;
    c->d_korder = d_order;
    c->korder_n = (int)order.size();
    c->net_q_fw = 0.0;
    for (int i = 0; i < c->n; ++i)
      if (considered_charged(c->q[i])) c->net_q_fw += c->q[i];
  }
  *out = c.release();
  return RHG_OK;
}
} // namespace

int rhg_create_on_stream(const rhg_framework *fw, const rhg_cell *cell, const rhg_forcefield *ff, int device,
                         hipStream_t stream, rhg_workspace *ws, rhg_pool *pool, rhg_ctx **out) {
  return create_impl(fw, cell, ff, device, stream, ws, pool, out);
}

int rhg_use_stream(rhg_ctx *c, hipStream_t stream) {
  if (!c) return RHG_ERR_INVALID;
  if (c->stream == stream) return RHG_OK;
  hipSetDevice(c->device);
  if (c->stream) {
    if (int rs_ = sync_stream(c)) return rs_;
  }
  if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
  c->stream = stream;
  c->own_stream = false;
  return RHG_OK;
}

extern "C" {

int rhg_create(const rhg_framework *fw, const rhg_cell *cellp, const rhg_forcefield *ff, int device,
               rhg_ctx **out) {
  return create_impl(fw, cellp, ff, device, nullptr, nullptr, nullptr, out);
}

void rhg_destroy(rhg_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->ws && c->ws->owner == c) c->ws->owner = nullptr;
  for (auto &b : c->bufs) {
    dev_free(c->arena.pool, b.second.p, b.second.bytes);
    for (auto &r : b.second.retired) dev_free(c->arena.pool, r.first, r.second);
  }
  arena_release(c); // every table allocated with dalloc
  for (auto &kv : c->graph_cache) hipGraphExecDestroy(kv.second);
  for (auto &p : c->pending) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) hipEventDestroy(e);
  // the main stream waited for the second one's last work (its join event)
  if (c->aux) {
    hipStreamSynchronize(c->aux);
    hipStreamDestroy(c->aux);
  }
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int rhg_add_pseudo_atom(rhg_ctx *c, double eps_k, double sigma, double charge, int32_t *type_out) {
  if (!c || !type_out) return RHG_ERR_INVALID;
  if (!(eps_k >= 0.0) || !(sigma > 0.0)) return fail(c, RHG_ERR_INVALID, "bad LJ parameters");
  hipSetDevice(c->device);
  PseudoAtom t;
  t.eps_k = eps_k;
  t.sigma = sigma;
  t.q = charge;
  int rc = setup_pseudo_atom(c, t);
  if (!rc) rc = ads_records(c, t);
  if (!rc) rc = cell_shift_room(c, t);
  if (rc) return rc;
  c->types.push_back(t);
  *type_out = (int32_t)c->types.size() - 1;
  return RHG_OK;
}

int rhg_set_precision(rhg_ctx *c, int32_t p) {
  if (!c) return RHG_ERR_INVALID;
  if (p != RHG_PREC_FP64 && p != RHG_PREC_MIXED) return fail(c, RHG_ERR_INVALID, "bad precision");
  c->precision = p;
  return RHG_OK;
}

int rhg_footprint(const rhg_ctx *c, size_t *bytes) {
  if (!c || !bytes) return RHG_ERR_INVALID;
  size_t b = c->footprint;
  for (auto &kv : c->bufs) b += kv.second.bytes;
  *bytes = b;
  return RHG_OK;
}

int rhg_info(const rhg_ctx *c, double out[10]) {
  if (!c || !out) return RHG_ERR_INVALID;
  out[0] = c->alpha;
  out[1] = c->kx;
  out[2] = c->ky;
  out[3] = c->kz;
  out[4] = c->nk;
  out[5] = c->uion;
  out[6] = c->beta;
  out[7] = c->props[6];
  out[8] = c->mass_total;
  // src/statistics.c:4751
  out[9] = 1e-3 * c->mass_total / (c->props[6] * (kAngstrom * kAngstrom * kAngstrom) * kAvogadro);
  return RHG_OK;
}

int rhg_framework_energy_at_positions(rhg_ctx *c, int64_t n, const double *xyz, int32_t type,
                                      double *u_vdw, double *u_cc, uint8_t *overlap) {
  if (!c || n < 0 || (n > 0 && !xyz)) return RHG_ERR_INVALID;
  if (int rc = check_type(c, type)) return rc;
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  std::vector<double> sx(n), sy(n), sz(n);
  for (int64_t i = 0; i < n; ++i) {
    sx[i] = xyz[3 * i];
    sy[i] = xyz[3 * i + 1];
    sz[i] = xyz[3 * i + 2];
  }
  size_t b = n * sizeof(double);
  double *dx = (double *)work(c, "e_x", b), *dy = (double *)work(c, "e_y", b), *dz = (double *)work(c, "e_z", b);
  double *dv = (double *)work(c, "e_v", b), *dc = (double *)work(c, "e_c", b);
  uint8_t *dov = (uint8_t *)work(c, "e_o", n);
  if (!dx || !dy || !dz || !dv || !dc || !dov) return fail(c, RHG_ERR_OOM, "work buffers");
  upload(c, dx, sx.data(), b);
  upload(c, dy, sy.data(), b);
  upload(c, dz, sz.data(), b);
  EnergyArgs a = energy_args(c, type, 1.0, 1, 1);
  if (!a.do_cc) fill(dc, n, 0.0, c->stream);
  SiteList s{dx, dy, dz, n, nullptr, 1};
  EnergyOut o{dv, a.do_cc ? dc : nullptr, dov};
  if (int rc = site_energy(c, type, a, s, o, "e")) return rc;
  if (!a.do_cc) {
    // overlap from the VDW term only; the charge term is zero
  }
  HIPC(c, hipGetLastError());
  if (u_vdw) HIPC(c, hipMemcpyAsync(u_vdw, dv, b, hipMemcpyDeviceToHost, c->stream));
  if (u_cc) HIPC(c, hipMemcpyAsync(u_cc, dc, b, hipMemcpyDeviceToHost, c->stream));
  if (overlap) HIPC(c, hipMemcpyAsync(overlap, dov, n, hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

int rhg_framework_energy_at_positions_noalloc(rhg_ctx *c, int64_t n, const double *d_xyz, int32_t type,
                                              double *d_u_vdw, double *d_u_cc, uint8_t *d_overlap) {
  if (!c || n < 0 || (n > 0 && !d_xyz)) return RHG_ERR_INVALID;
  if (int rc = check_type(c, type)) return rc;
  if (n > kNaSites)
    return fail(c, RHG_ERR_INVALID, "noalloc: at most 262144 positions per call");
  if (n == 0) return RHG_OK;
  if (c->device < 0 || c->device >= kNaDevices) return fail(c, RHG_ERR_INVALID, "noalloc: device ordinal >= 64");
  unsigned char *base = na_ptr[c->device];
  if (!base)
    return fail(c, RHG_ERR_HIP, std::string("noalloc: hipGetSymbolAddress(scratch): ") +
                                    hipGetErrorString(na_err[c->device]));
  const bool culled = c->cull && c->n > 0;
  if (culled && (!c->grid_ready || (c->precision != RHG_PREC_FP64 && c->use_images && !c->img_ready)))
    return fail(c, RHG_ERR_INVALID,
                "noalloc: the context's cell lists are not built yet (one call of "
                "rhg_framework_energy_at_positions at set-up builds them)");
  if (culled && c->cg.ncell > kNaCells) return fail(c, RHG_ERR_INVALID, "noalloc: cell grid larger than 64^3");
  hipSetDevice(c->device);
  // carve the array; the per-cell counters are cleared by the binning itself
  unsigned char *p = base;
  double *sx = (double *)p;
  p += kNaXyz;
  double *sy = (double *)p;
  p += kNaXyz;
  double *sz = (double *)p;
  p += kNaXyz;
  BinWork b;
  b.cell_of = (int32_t *)p;
  p += kNaI32;
  b.perm = (int32_t *)p;
  p += kNaI32;
  b.fix = (float4 *)p;
  p += kNaFix;
  b.count = (int32_t *)p;
  p += kNaCell;
  b.rank = (int32_t *)p; // kNaSites ranks in a per-cell slot (kNaCell >= kNaI32)
  p += kNaCell;
  b.start = (int32_t *)p;
  p += kNaCell;
  b.tile_off = (int32_t *)p;
  p += kNaCell;
  b.redo = (int2 *)p;
  p += 2 * kNaI32;
  b.n_redo = (int32_t *)p;
  b.pairs = nullptr; // no work counters on this path
  // event timing creates events: off for the duration of the call
  const int prof = c->prof;
  c->prof = 0;
  launch_deinterleave(d_xyz, n, sx, sy, sz, c->stream);
  EnergyArgs a = energy_args(c, type, 1.0, 1, 1);
  if (!a.do_cc && d_u_cc) fill(d_u_cc, n, 0.0, c->stream);
  SiteList s{sx, sy, sz, n, nullptr, 1};
  EnergyOut o{d_u_vdw, a.do_cc ? d_u_cc : nullptr, d_overlap};
  int rc = site_energy(c, type, a, s, o, "e", culled ? &b : nullptr);
  c->prof = prof;
  if (rc) return rc;
  HIPC(c, hipGetLastError());
  return RHG_OK;
}

int rhg_framework_vdw_energy_at_position(rhg_ctx *c, const double pos[3], int32_t type, double scaling,
                                         double *u) {
  (void)scaling; // not used by the LENNARD_JONES branch
  if (!c || !pos || !u) return RHG_ERR_INVALID;
  return rhg_framework_energy_at_positions(c, 1, pos, type, u, nullptr, nullptr);
}

int rhg_framework_charge_energy_at_position(rhg_ctx *c, const double pos[3], int32_t type,
                                            double scaling, double *ucc, double *ucbd) {
  if (!c || !pos || !ucc) return RHG_ERR_INVALID;
  if (int rc = check_type(c, type)) return rc;
  if (ucbd) *ucbd = 0.0;
  *ucc = 0.0;
  EnergyArgs a = energy_args(c, type, scaling, 0, 1);
  if (!a.do_cc) return RHG_OK;
  hipSetDevice(c->device);
  double *d = (double *)work(c, "e_x", 4 * sizeof(double));
  double *dv = (double *)work(c, "e_c", sizeof(double));
  if (!d || !dv) return RHG_ERR_OOM;
  upload(c, d, pos, 3 * sizeof(double));
  SiteList s{d, d + 1, d + 2, 1, nullptr, 1};
  EnergyOut o{nullptr, dv, nullptr};
  if (c->coulomb_grid) {
    if (int rc = site_energy(c, type, a, s, o, "e")) return rc;
  } else {
    launch_site_energy(a, s, o, c->stream);
  }
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(ucc, dv, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

int rhg_ewald_tables(rhg_ctx *c, double *kfactor, double *sf_re, double *sf_im) {
  if (!c) return RHG_ERR_INVALID;
  if (c->charge_method != RHG_CHARGE_EWALD) return fail(c, RHG_ERR_UNSUPPORTED, "no Ewald");
  hipSetDevice(c->device);
  if (kfactor) memcpy(kfactor, c->kfactor.data(), c->nk * sizeof(double));
  if (sf_re || sf_im) {
    if (int rc = ensure_fourier(c)) return rc;
    std::vector<double2> h(c->nk);
    HIPC(c, hipMemcpyAsync(h.data(), c->d_sf, c->nk * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    if (int rs_ = sync_stream(c)) return rs_;
    for (int k = 0; k < c->nk; ++k) {
      if (sf_re) sf_re[k] = h[k].x;
      if (sf_im) sf_im[k] = h[k].y;
    }
  }
  return RHG_OK;
}

} // extern "C"

namespace {

int build_mol(rhg_ctx *c, const rhg_molecule *m, MolDev &d) {
  if (!m || m->n_sites < 1 || m->n_sites > RHG_MAX_SITES) return fail(c, RHG_ERR_INVALID, "bad molecule");
  if (m->start < 0 || m->start >= m->n_sites) return fail(c, RHG_ERR_INVALID, "bad starting bead");
  if (m->k_fb < 1 || m->k_fb > kMaxTrials) return fail(c, RHG_ERR_INVALID, "k_fb out of range [1,64]");
  if (m->n_sites > 1 && (m->k_or < 1 || m->k_or > kMaxTrials))
    return fail(c, RHG_ERR_INVALID, "k_or out of range [1,64]");
  memset(&d, 0, sizeof(d));
  d.n_sites = m->n_sites;
  d.start = m->start;
  d.k_fb = m->k_fb;
  d.k_or = m->n_sites > 1 ? m->k_or : 1;
  d.n_place = 0;
  for (int j = 0; j < m->n_sites; ++j) {
    if (m->type[j] < 0 || m->type[j] >= (int)c->types.size()) return fail(c, RHG_ERR_INVALID, "unknown site type");
    for (int q = 0; q < 3; ++q) d.body[j][q] = m->body[j][q];
    d.q[j] = c->types[m->type[j]].q;
    d.charged[j] = considered_charged(d.q[j]) ? 1 : 0;
    if (j != m->start) d.place[d.n_place++] = j;
  }
  // storage groups: placed beads whose pseudo-atoms have the same
  // parameters are stored (and evaluated) together; groups in order of their
  // first bead
  auto same_params = [&](int ta, int tb) {
    const PseudoAtom &A = c->types[ta], &B = c->types[tb];
    return ta == tb || (A.eps_k == B.eps_k && A.sigma == B.sigma && A.q == B.q &&
                        vdw_grid_of(c, ta) == vdw_grid_of(c, tb));
  };
  auto group = [&](bool charged_only, int *start, int *size, int *b) {
    int next = 0;
    bool done[kMaxSites] = {false};
    for (int p = 0; p < d.n_place; ++p) {
      start[p] = -1;
      size[p] = 0;
      b[p] = 0;
    }
    for (int p = 0; p < d.n_place; ++p) {
      if (done[p] || (charged_only && !d.charged[d.place[p]])) continue;
      int cnt = 0;
      for (int q = p; q < d.n_place; ++q)
        if (!done[q] && same_params(m->type[d.place[q]], m->type[d.place[p]]) &&
            !(charged_only && !d.charged[d.place[q]])) {
          start[q] = next;
          b[q] = cnt++;
          done[q] = true;
        }
      for (int q = p; q < d.n_place; ++q)
        if (start[q] == next) size[q] = cnt;
      next += cnt;
    }
  };
  group(false, d.g_start, d.g_size, d.g_b);
  group(true, d.h_start, d.h_size, d.h_b);
  if (c->charge_method != RHG_CHARGE_EWALD)
    for (int p = 0; p < d.n_place; ++p) d.h_start[p] = -1;
  return RHG_OK;
}

struct WidomBufs {
  WidomWork w;
  double *d_part;
};

constexpr int kSeg = 4096; // insertions per block-reduction segment

int widom_buffers(rhg_ctx *c, const MolDev &m, int64_t cap, WidomBufs &B) {
  WidomWork &w = B.w;
  memset(&w, 0, sizeof(w));
  w.cap = cap;
  size_t nfb = (size_t)cap * m.k_fb;
  size_t nor = (size_t)cap * m.k_or * std::max(m.n_place, 1);
  size_t ncf = (size_t)cap * m.n_sites;
  // site lists are indexed with 32-bit positions
  if (std::max(nfb, std::max(nor, ncf)) > (size_t)INT32_MAX)
    return fail(c, RHG_ERR_INVALID, "chunk x trials x sites above 2^31 - 1: use a smaller chunk");
  size_t d8 = sizeof(double);
#define WB(field, name, bytes)                           \
  w.field = (decltype(w.field))work(c, name, (bytes)); \
  if (!w.field) return fail(c, RHG_ERR_OOM, "widom work " name);
  WB(fbx, "w_fbx", nfb * d8);
  WB(fby, "w_fby", nfb * d8);
  WB(fbz, "w_fbz", nfb * d8);
  WB(fb_uv, "w_fbuv", nfb * d8);
  WB(fb_uc, "w_fbuc", nfb * d8);
  WB(stage, "w_stage", cap * 4);
  WB(selfb, "w_selfb", cap * 4);
  WB(selor, "w_selor", cap * 4);
  WB(fbmask, "w_fbmask", cap * 8);
  WB(ormask, "w_ormask", cap * 8);
  WB(W1, "w_W1", cap * d8);
  WB(W2, "w_W2", cap * d8);
  WB(W, "w_W", cap * d8);
  WB(dU, "w_dU", cap * d8);
  WB(uvdw, "w_uvdw", cap * d8);
  WB(ureal, "w_ureal", cap * d8);
  WB(ufh, "w_ufh", cap * d8);
  WB(ufa, "w_ufa", cap * d8);
  WB(usel_or, "w_uselor", cap * d8);
  WB(selx, "w_selx", cap * d8);
  WB(sely, "w_sely", cap * d8);
  WB(selz, "w_selz", cap * d8);
  WB(active, "w_active", cap * 4);
  WB(n_active, "w_nactive", 64);
  WB(orx, "w_orx", nor * d8);
  WB(ory, "w_ory", nor * d8);
  WB(orz, "w_orz", nor * d8);
  WB(or_uv, "w_oruv", nor * d8);
  WB(active2, "w_active2", cap * 4);
  WB(n_active2, "w_nactive2", 64);
  WB(cfx, "w_cfx", ncf * d8);
  WB(cfy, "w_cfy", ncf * d8);
  WB(cfz, "w_cfz", ncf * d8);
  WB(cc_place, "w_ccplace", (size_t)cap * std::max(m.n_place, 1) * d8);
  WB(ccx, "w_ccx", (size_t)cap * std::max(m.n_place, 1) * d8);
  WB(ccy, "w_ccy", (size_t)cap * std::max(m.n_place, 1) * d8);
  WB(ccz, "w_ccz", (size_t)cap * std::max(m.n_place, 1) * d8);
  WB(f_host, "w_fhost", cap * d8);
  WB(f_ads, "w_fads", cap * d8);
  if (c->ads_n > 0) {
    WB(fb_av, "w_fbav", nfb * d8);
    WB(fb_ac, "w_fbac", nfb * d8);
    WB(or_av, "w_orav", nor * d8);
    WB(cc_ads, "w_ccads", (size_t)cap * std::max(m.n_place, 1) * d8);
    WB(uadsv, "w_uadsv", cap * d8);
    WB(uadsc, "w_uadsc", cap * d8);
  }
#undef WB
  B.d_part = (double *)work(c, "w_part", ((cap + kSeg - 1) / kSeg + 1) * 10 * d8);
  if (!B.d_part) return fail(c, RHG_ERR_OOM, "widom reduce buffers");
  return RHG_OK;
}

// accumulators and control block of the context (allocated once)
int widom_state(rhg_ctx *c) {
  if (c->d_acc) return RHG_OK;
  int rc;
  if ((rc = dalloc(c, (void **)&c->d_acc, 32 * sizeof(double)))) return rc;
  if ((rc = dalloc(c, (void **)&c->d_ctl, CTL_N * sizeof(int64_t)))) return rc;
  if ((rc = dalloc(c, (void **)&c->d_neff, 64))) return rc;
  fill(c->d_acc, 32, 0.0, c->stream);
  c->gen++;
  return RHG_OK;
}

StatArgs stat_args(rhg_ctx *c, const rhg_molecule *mol, const MolDev &m, int ewald) {
  StatArgs sa;
  sa.n_atoms = (double)c->n;
  sa.nk = (double)c->nk;
  sa.k_fb = m.k_fb;
  sa.k_or = m.k_or;
  sa.n_place = m.n_place;
  sa.n_charged = 0;
  for (int j = 0; j < m.n_sites; ++j) sa.n_charged += m.charged[j];
  sa.n_cc_place = 0;
  if (m.n_sites > 1 && c->charge_method == RHG_CHARGE_EWALD)
    for (int p = 0; p < m.n_place; ++p) sa.n_cc_place += m.charged[m.place[p]];
  const int type0 = mol->type[m.start];
  sa.fb_charged = (c->charge_method == RHG_CHARGE_EWALD && considered_charged(c->types[type0].q)) ? 1 : 0;
  sa.ewald = ewald;
  sa.multi = m.n_sites > 1 ? 1 : 0;
  return sa;
}

// One chunk of the Widom pipeline, all launches on the context stream and
// none of them waiting for the host: trial generation, first-bead energies
// and selection, orientation energies and selection, Coulomb of the
// selected beads, Fourier change, finalisation, then (accumulate) the block
// sums and work counters into the device accumulators and the control block
// moved on by the chunk.
// Starts work on the context's second stream after everything enqueued on
// its main stream so far; *forked is false (nothing done) for a borrowed
// stream (a screening slot, whose streams already fill the hardware queues),
// with every kernel family timed (the events are on the main stream) and
// while the main stream is being captured into a graph.
int fork_aux(rhg_ctx *c, bool *forked) {
  *forked = false;
  if (!c->own_stream || c->prof == 1) return RHG_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIPC(c, hipStreamIsCapturing(c->stream, &cs));
  if (cs != hipStreamCaptureStatusNone) return RHG_OK;
  if (!c->aux) {
    HIPC(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    HIPC(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    HIPC(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  }
  HIPC(c, hipEventRecord(c->ev_fork, c->stream));
  HIPC(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
  *forked = true;
  return RHG_OK;
}

int widom_chunk(rhg_ctx *c, const rhg_molecule *mol, const MolDev &m, WidomBufs &B,
                const WidomArgs &a, bool accumulate) {
  WidomWork &w = B.w;
  const int64_t n = a.n;
  hipStream_t st = c->stream;
  {
    // the generator also clears the selection kernels' slot counters
    Timed t(c, "widom_gen");
    launch_widom_fb_gen(a, w, st);
  }
  // HandleFirstBead energies: VDW and real-space Coulomb of every trial
  // (src/cbmc.c:680-681)
  const int type0 = mol->type[m.start];
  {
    EnergyArgs ea = energy_args(c, type0, 1.0, 1, 1);
    if (!ea.do_cc) fill(w.fb_uc, w.cap * m.k_fb, 0.0, st);
    // trials stored [trial][insertion]
    SiteList s{w.fbx, w.fby, w.fbz, n * m.k_fb, a.n_eff, m.k_fb, w.cap};
    EnergyOut o{w.fb_uv, ea.do_cc ? w.fb_uc : nullptr, nullptr};
    if (int rc = site_energy(c, type0, ea, s, o, "bfb")) return rc;
    if (w.fb_av) {
      // guest-guest terms; the inserted molecule is number ads_nmol, so none
      // is excluded (src/mc_moves.c:7856, src/cbmc.c:683-689)
      AdsArgs aa = ads_args(c, type0, a.exclude, 1, 1);
      aa.excl = a.excl; // columns of the first-bead rows are insertions
      if (!aa.do_cc) fill(w.fb_ac, w.cap * m.k_fb, 0.0, st);
      launch_site_ads(aa, s, w.fb_av, aa.do_cc ? w.fb_ac : nullptr, st);
    }
  }
  {
    Timed t(c, "widom_select");
    launch_widom_fb_select(a, w, st);
  }
  // Fourier Delta of the accepted molecules (CalculateEwaldFourierAdsorbate)
  // on stream fs; it needs only the selected configuration, so for
  // multi-site molecules it can run beside the placed-bead Coulomb stage
  auto fourier_stage = [&](hipStream_t fs) -> int {
    if (!a.ewald) return RHG_OK;
    FourierArgs f;
    memset(&f, 0, sizeof(f));
    f.e = ewald_dev(c);
    f.n_sites = m.n_sites;
    f.nc = 0;
    f.self_energy = 0.0;
    for (int j = 0; j < m.n_sites; ++j)
      if (m.charged[j]) {
        f.cidx[f.nc] = j;
        f.q[f.nc] = 1.0 * m.q[j];
        // src/ewald.c:6973
        f.self_energy += c->u.C * (f.q[f.nc] * f.q[f.nc]) * c->alpha / sqrt(M_PI);
        f.nc++;
      }
    f.px = w.cfx;
    f.py = w.cfy;
    f.pz = w.cfz;
    f.stride = w.cap;
    f.n = n;
    f.n_dev = w.n_active2;
    f.d_host = w.f_host;
    f.d_ads = w.f_ads;
    f.sads = c->ads_n > 0 ? c->d_sads : nullptr;
    f.net_q_ads = c->net_q_ads;
    if (f.nc > 0) {
      double2 *part = (double2 *)work(c, "w_fpart", (size_t)(c->kx + 1) * w.cap * sizeof(double2));
      if (!part) return fail(c, RHG_ERR_OOM, "fourier partials");
      Timed t(c, "fourier");
      launch_fourier(f, part, fs);
    } else {
      fill(w.f_host, n, 0.0, fs);
      fill(w.f_ads, n, 0.0, fs);
    }
    return RHG_OK;
  };
  bool forked = false;
  if (m.n_sites > 1) {
    // ComputeExternalEnergies VDW of every orientation trial (src/cbmc.c:1210),
    // one site list per type group of placed beads; positions and energies
    // are stored [bead][orientation][active slot]
    const int64_t cap_or = w.cap * (int64_t)m.k_or;
    for (int p = 0; p < m.n_place; ++p) {
      if (m.g_b[p] != 0) continue; // first bead of its group
      EnergyArgs ea = energy_args(c, mol->type[m.place[p]], 1.0, 1, 0);
      const int64_t base = m.g_start[p] * cap_or;
      const int rows = m.g_size[p] * m.k_or;
      SiteList s{w.orx + base, w.ory + base, w.orz + base, n * rows, w.n_active, rows, w.cap};
      EnergyOut o{w.or_uv + base, nullptr, nullptr};
      if (int rc = site_energy(c, mol->type[m.place[p]], ea, s, o, "bor")) return rc;
      if (w.or_av) {
        AdsArgs aa = ads_args(c, mol->type[m.place[p]], a.exclude, 1, 0);
        aa.excl = a.excl;
        aa.col_map = w.active; // columns are active slots
        launch_site_ads(aa, s, w.or_av + base, nullptr, st);
      }
    }
    {
      Timed t(c, "widom_select");
      launch_widom_or_select(a, w, st);
    }
    if (a.ewald) {
      // the structure factor (first use) on the main stream, before the fork
      for (int j = 0; j < m.n_sites; ++j)
        if (m.charged[j]) {
          if (int rc = ensure_fourier(c)) return rc;
          break;
        }
      hipStream_t fs = st;
      if (int rc = fork_aux(c, &forked)) return rc;
      if (forked) fs = c->aux;
      if (int rc = fourier_stage(fs)) return rc;
      if (forked) HIPC(c, hipEventRecord(c->ev_join, c->aux));
    }
    // real-space Coulomb of the charged placed beads of the selected
    // orientation (src/cbmc.c:1216; only the selected trial's value is used),
    // one site list per type group
    for (int p = 0; p < m.n_place; ++p) {
      if (m.h_start[p] < 0 || m.h_b[p] != 0) continue;
      EnergyArgs ea = energy_args(c, mol->type[m.place[p]], 1.0, 0, 1);
      const int64_t base = m.h_start[p] * w.cap;
      const int rows = m.h_size[p];
      if (!ea.do_cc) {
        fill(w.cc_place + base, (int64_t)rows * w.cap, 0.0, st);
        if (w.cc_ads) fill(w.cc_ads + base, (int64_t)rows * w.cap, 0.0, st);
        continue;
      }
      SiteList s{w.ccx + base, w.ccy + base, w.ccz + base, n * rows, w.n_active2, rows, w.cap};
      EnergyOut o{nullptr, w.cc_place + base, nullptr};
      if (int rc = site_energy(c, mol->type[m.place[p]], ea, s, o, "bcc")) return rc;
      if (w.cc_ads) {
        AdsArgs aa = ads_args(c, mol->type[m.place[p]], a.exclude, 0, 1);
        aa.excl = a.excl;
        aa.col_map = w.active2; // columns are accepted slots
        launch_site_ads(aa, s, nullptr, w.cc_ads + base, st);
      }
    }
  }
  if (m.n_sites <= 1) {
    for (int j = 0; j < m.n_sites; ++j)
      if (a.ewald && m.charged[j])
        if (int rc = ensure_fourier(c)) return rc;
    if (int rc = fourier_stage(st)) return rc;
  }
  if (forked) HIPC(c, hipStreamWaitEvent(st, c->ev_join, 0));
  {
    Timed t(c, "widom_select");
    launch_widom_finalize(a, w, st);
  }
  if (accumulate) {
    Timed t(c, "block_reduce");
    launch_chunk_tail(a, w, stat_args(c, mol, m, a.ewald), w.W, w.dU, B.d_part, kSeg, c->d_acc,
                      c->d_acc + 16, c->d_ctl, n, st);
  } else {
    launch_advance(c->d_ctl, n, st);
  }
  if (hipGetLastError() != hipSuccess) return fail(c, RHG_ERR_HIP, "widom launch failed");
  return RHG_OK;
}

// block boundaries (insertion index) of a job (src/statistics.c:898-910,1126)
void block_bounds(int64_t n_total, int spc, int64_t bounds[6]) {
  int64_t n_cycles = (n_total + spc - 1) / spc;
  long long BlockCycle[RHG_NR_BLOCKS];
  for (int i = 0; i < RHG_NR_BLOCKS; i++)
    BlockCycle[i] = (long long)((i + 1) * n_cycles / (double)RHG_NR_BLOCKS);
  // first cycle of each block
  int64_t first_cycle[RHG_NR_BLOCKS + 1];
  for (int b = 0; b <= RHG_NR_BLOCKS; ++b) first_cycle[b] = n_cycles;
  first_cycle[0] = 0;
  int Block = 0;
  for (int64_t cyc = 0; cyc < n_cycles; cyc++) {
    if (Block < RHG_NR_BLOCKS && cyc == BlockCycle[Block]) {
      Block++;
      if (Block < RHG_NR_BLOCKS) first_cycle[Block] = cyc;
    }
  }
  // blocks that are never entered start at the end
  for (int b = 1; b < RHG_NR_BLOCKS; ++b)
    if (first_cycle[b] < first_cycle[b - 1]) first_cycle[b] = first_cycle[b - 1];
  for (int b = 0; b < RHG_NR_BLOCKS; ++b) bounds[b] = std::min(first_cycle[b] * spc, n_total);
  bounds[RHG_NR_BLOCKS] = n_total;
}

// WidomIdealGasAccumulated of insertions [first, first+n) of a job whose
// blocks start at jb: weight x Volume / (NumberOfMolecules + 1) per insertion
// (src/mc_moves.c:7861-7863; CFBiasingWeight is 1 on this path), with the
// integer molecules present now (the adsorbates; no cations on this path).
void note_ideal_gas(rhg_ctx *c, const int64_t jb[6], int64_t first, int64_t n) {
  const double per = c->props[6] / ((double)c->ads_nmol + 1.0);
  for (int b = 0; b < RHG_NR_BLOCKS; ++b) {
    const int64_t lo = std::max(first, jb[b]), hi = std::min(first + n, jb[b + 1]);
    if (hi > lo) c->ig_pending[b] += (double)(hi - lo) * per;
  }
}

WidomArgs widom_args(rhg_ctx *c, const MolDev &m, uint64_t seed, int64_t cap) {
  WidomArgs a;
  memset(&a, 0, sizeof(a));
  a.pockets = c->pockets;
  a.mol = m;
  a.cell = c->cell;
  a.beta = c->beta;
  a.seed = seed;
  a.n = cap;
  a.ctl = c->d_ctl;
  a.n_eff = c->d_neff;
  a.use_host_trials = 0;
  a.ewald = (c->charge_method == RHG_CHARGE_EWALD) ? 1 : 0;
  a.retrace = 0;
  a.exclude = c->ads_nmol; // an inserted molecule is number ads_nmol: none left out
  return a;
}

int check_fault(rhg_ctx *c) {
  if (c->fault_after < 0) return RHG_OK;
  if (c->fault_after == 0) {
    c->fault_after = -1;
    return fail(c, RHG_ERR_FAULT_INJECTED, "fault injected by the host-side hook");
  }
  c->fault_after--;
  return RHG_OK;
}

// Enqueues insertions [first, first+n) of a job of n_total insertions; the
// block sums and work counters go into the device accumulators.  With
// graphs enabled and more than one chunk, the first chunk runs eagerly (it
// sizes every buffer) and the others replay one captured chunk.
int widom_enqueue(rhg_ctx *c, const rhg_molecule *mol, const MolDev &m, uint64_t seed, int64_t n_total,
                  int64_t first, int64_t n, int32_t spc, double *rec) {
  hipStream_t st = c->stream;
  const int64_t cap = std::min<int64_t>(n, c->chunk);
  WidomBufs B;
  if (int rc = widom_state(c)) return rc;
  if (int rc = widom_buffers(c, m, cap, B)) return rc;
  int64_t jb[6];
  block_bounds(n_total, spc, jb);
  launch_set_ctl(c->d_ctl, first, first + n, jb, st);
  const int64_t n_chunks = (n + cap - 1) / cap;
  const bool use_graph = c->graphs && !c->prof && !rec && n_chunks > 1 && c->fault_after < 0;
  for (int64_t k = 0; k < n_chunks; ++k) {
    const int64_t off = k * cap;
    const int64_t cnt = std::min(cap, n - off);
    if (int rc = check_fault(c)) return rc;
    note_ideal_gas(c, jb, first + off, cnt);
    if (use_graph && k > 0) {
      // one captured chunk, keyed by everything its launches hold
      std::string key((const char *)&m, sizeof(m));
      key.append((const char *)mol->type, sizeof(mol->type));
      key.append((const char *)&seed, sizeof(seed));
      key.append((const char *)&cap, sizeof(cap));
      key.append((const char *)&c->gen, sizeof(c->gen));
      const uint64_t wgen = c->ws ? c->ws->gen : 0, wuid = c->ws ? c->ws->uid : 0;
      key.append((const char *)&wuid, sizeof(wuid));
      key.append((const char *)&wgen, sizeof(wgen));
      key.append((const char *)&c->precision, sizeof(c->precision));
      key.append((const char *)&c->cull, sizeof(c->cull));
      auto it = c->graph_cache.find(key);
      if (it == c->graph_cache.end()) {
        const uint64_t gen0 = c->gen, wgen0 = c->ws ? c->ws->gen : 0;
        WidomArgs a = widom_args(c, m, seed, cap);
        HIPC(c, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        int rc = widom_chunk(c, mol, m, B, a, true);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) {
          if (g) hipGraphDestroy(g);
          return rc;
        }
        if (e != hipSuccess) return fail(c, RHG_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(c, RHG_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(e));
        if (c->gen != gen0 || (c->ws ? c->ws->gen : 0) != wgen0) {
          // a buffer moved while capturing: this graph is stale
          hipGraphExecDestroy(ge);
          return fail(c, RHG_ERR_HIP, "buffers changed during capture");
        }
        if (c->graph_cache.size() >= 16) {
          for (auto &kv : c->graph_cache) hipGraphExecDestroy(kv.second);
          c->graph_cache.clear();
        }
        it = c->graph_cache.emplace(key, ge).first;
      }
      HIPC(c, hipGraphLaunch(it->second, st));
      continue;
    }
    WidomArgs a = widom_args(c, m, seed, use_graph ? cap : cnt);
    if (int rc = widom_chunk(c, mol, m, B, a, true)) return rc;
    if (rec) {
      double *drec = (double *)work(c, "w_rec", (size_t)cnt * RHG_REC_N * sizeof(double));
      if (!drec) return fail(c, RHG_ERR_OOM, "rec buffer");
      launch_widom_records(a, B.w, drec, st);
      HIPC(c, hipMemcpyAsync(rec + off * RHG_REC_N, drec, (size_t)cnt * RHG_REC_N * sizeof(double),
                             hipMemcpyDeviceToHost, st));
    }
  }
  return RHG_OK;
}

// Adds the device block sums to `blocks` and clears them (waits for the
// stream).
int widom_collect(rhg_ctx *c, rhg_widom_blocks *blocks) {
  if (!c->d_acc) return RHG_OK;
  double acc[16];
  HIPC(c, hipMemcpyAsync(acc, c->d_acc, 15 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  fill(c->d_acc, 15, 0.0, c->stream);
  double ig[RHG_NR_BLOCKS];
  for (int b = 0; b < RHG_NR_BLOCKS; ++b) {
    ig[b] = c->ig_pending[b];
    c->ig_pending[b] = 0.0;
  }
  if (int rs_ = sync_stream(c)) return rs_;
  if (blocks) {
    for (int b = 0; b < RHG_NR_BLOCKS; ++b) {
      blocks->count[b] += acc[10 + b];
      blocks->ideal_gas[b] += ig[b];
      blocks->sum_w[b] += acc[b];
      blocks->sum_wdu[b] += acc[5 + b];
    }
  }
  return RHG_OK;
}

int check_range(rhg_ctx *c, const rhg_molecule *mol, int64_t n_total, int64_t first, int64_t n,
                int32_t spc) {
  if (!c || !mol) return RHG_ERR_INVALID;
  if (n_total < 0 || first < 0 || n < 0 || first + n > n_total || spc < 1)
    return fail(c, RHG_ERR_INVALID, "bad insertion range");
  return RHG_OK;
}

// ---------------------------------------------------------- flexible chains
// Growth plan of a molecule of one flexible group: SetConnectivityMatrix,
// SetGrowingStatus and Interactions (src/cbmc.c:2582-2760, 901-1152) run
// once on the host, since the path (always the first possible bead,
// src/cbmc.c:2651) does not depend on positions.  Each step becomes a
// ChainStep table with atom references resolved into local codes.
int chain_plan(rhg_ctx *c, const rhg_chain *ch, std::vector<ChainStep> &steps) {
  const int n = ch->n_atoms;
  if (n < 2 || n > kChainMax || ch->start < 0 || ch->start >= n || !ch->type)
    return fail(c, RHG_ERR_INVALID, "chain: n_atoms out of [2, 64] or bad starting bead");
  if (ch->k_fb < 1 || ch->k_fb > kMaxTrials || ch->k_trial < 1 || ch->k_trial > kMaxTrials ||
      ch->k_torsion < ch->k_trial || ch->k_torsion > kChainTorsionTrials || ch->moves_per_open_bead < 0)
    return fail(c, RHG_ERR_INVALID,
                "chain: k_fb and k_trial must be in [1, 64], k_torsion in [k_trial, 1024], moves >= 0");
  for (int a = 0; a < n; ++a) {
    if (ch->type[a] < 0 || ch->type[a] >= (int)c->types.size()) return fail(c, RHG_ERR_INVALID, "chain: unknown pseudo-atom");
    if (considered_charged(c->types[ch->type[a]].q))
      return fail(c, RHG_ERR_UNSUPPORTED, "chain: charged beads are not supported on the flexible-chain path");
  }
  auto bad_atom = [&](int a) { return a < 0 || a >= n; };
  if (ch->n_bonds < 0 || ch->n_bends < 0 || ch->n_torsions < 0 || ch->n_intra_vdw < 0 ||
      (ch->n_bonds && (!ch->bonds || !ch->bond_par)) || (ch->n_bends && (!ch->bends || !ch->bend_par)) ||
      (ch->n_torsions && (!ch->torsions || !ch->torsion_par)) || (ch->n_intra_vdw && !ch->intra_vdw))
    return fail(c, RHG_ERR_INVALID, "chain: missing bonded-term arrays");
  for (int j = 0; j < ch->n_bonds; ++j) {
    const int32_t *b = ch->bonds + 3 * j;
    if (bad_atom(b[0]) || bad_atom(b[1]) || b[0] == b[1]) return fail(c, RHG_ERR_INVALID, "chain: bad bond");
    if (b[2] != RHG_BOND_FIXED && b[2] != RHG_BOND_HARMONIC) return fail(c, RHG_ERR_UNSUPPORTED, "chain: bond type");
    if (b[2] == RHG_BOND_HARMONIC && !(ch->bond_par[4 * j] > 0.0)) return fail(c, RHG_ERR_INVALID, "chain: bond constant");
  }
  for (int j = 0; j < ch->n_bends; ++j) {
    const int32_t *b = ch->bends + 4 * j;
    if (bad_atom(b[0]) || bad_atom(b[1]) || bad_atom(b[2])) return fail(c, RHG_ERR_INVALID, "chain: bad bend");
    if (b[3] != RHG_BEND_HARMONIC) return fail(c, RHG_ERR_UNSUPPORTED, "chain: bend type");
  }
  for (int j = 0; j < ch->n_torsions; ++j) {
    const int32_t *t = ch->torsions + 5 * j;
    if (bad_atom(t[0]) || bad_atom(t[1]) || bad_atom(t[2]) || bad_atom(t[3])) return fail(c, RHG_ERR_INVALID, "chain: bad torsion");
    if (t[4] != RHG_TORSION_TRAPPE) return fail(c, RHG_ERR_UNSUPPORTED, "chain: torsion type");
  }
  for (int j = 0; j < ch->n_intra_vdw; ++j)
    if (bad_atom(ch->intra_vdw[2 * j]) || bad_atom(ch->intra_vdw[2 * j + 1])) return fail(c, RHG_ERR_INVALID, "chain: bad intra pair");
  const double E2K = c->u.E2K;
  // intra-pair scaling (src/molecule.c:2997-3025)
  std::vector<double> scal((size_t)ch->n_intra_vdw, 1.0);
  for (int i = 0; i < ch->n_intra_vdw; ++i) {
    const int A = ch->intra_vdw[2 * i], B = ch->intra_vdw[2 * i + 1];
    for (int j = 0; j < ch->n_bends; ++j) {
      const int A1 = ch->bends[4 * j], B1 = ch->bends[4 * j + 1], C1 = ch->bends[4 * j + 2];
      for (int k = 0; k < ch->n_bends; ++k) {
        const int A2 = ch->bends[4 * k], B2 = ch->bends[4 * k + 1], C2 = ch->bends[4 * k + 2];
        if ((B1 == A2) && (C1 == B2) && (((A1 == A) && (C2 == B)) || ((A1 == B) && (C2 == A)))) scal[i] = ch->intra14_vdw_scaling;
        if ((B1 == C2) && (C1 == B2) && (((A1 == A) && (A2 == B)) || ((A1 == B) && (A2 == A)))) scal[i] = ch->intra14_vdw_scaling;
        if ((A1 == B2) && (B1 == A2) && (((C1 == A) && (C2 == B)) || ((C1 == B) && (C2 == A)))) scal[i] = ch->intra14_vdw_scaling;
        if ((A1 == B2) && (B1 == C2) && (((C1 == A) && (A2 == B)) || ((C1 == B) && (A2 == A)))) scal[i] = ch->intra14_vdw_scaling;
      }
    }
  }
  auto ref_len = [&](int A, int B) {
    for (int i = 0; i < ch->n_bonds; ++i) {
      const int32_t *b = ch->bonds + 3 * i;
      if ((b[0] == A && b[1] == B) || (b[1] == A && b[0] == B))
        return b[2] == RHG_BOND_HARMONIC ? ch->bond_par[4 * i + 1] : ch->bond_par[4 * i];
    }
    return 1.5;
  };
  auto per_atom = [&](const double *v, int a) { return v ? v[a] : 0.3; };
  std::vector<std::vector<char>> conn(n, std::vector<char>(n, 0));
  for (int j = 0; j < ch->n_bonds; ++j) {
    conn[ch->bonds[3 * j]][ch->bonds[3 * j + 1]] = 1;
    conn[ch->bonds[3 * j + 1]][ch->bonds[3 * j]] = 1;
  }
  auto todo = conn;
  std::vector<int> placed{ch->start};
  for (int j = 0; j < n; ++j) todo[ch->start][j] = 0;
  steps.clear();
  while ((int)placed.size() != n) {
    int beadn = -1, cur = -1;
    for (size_t i = 0; i < placed.size() && beadn < 0; ++i)
      for (int j = 0; j < n; ++j)
        if (todo[j][placed[i]]) {
          beadn = j;
          cur = placed[i];
          break;
        }
    if (beadn < 0) return fail(c, RHG_ERR_INVALID, "chain: the bonds do not connect every atom");
    ChainStep S;
    memset(&S, 0, sizeof(S));
    S.cur = cur;
    S.prev = -1;
    bool attached = false;
    for (int i = 0; i < n; ++i)
      if (conn[i][cur]) {
        if (todo[i][cur])
          attached = true;
        else {
          S.n_prev++;
          S.prev = i;
        }
      }
    std::vector<int> place;
    if (attached)
      for (int j = 0; j < n; ++j)
        if (todo[j][cur]) place.push_back(j);
    if (S.n_prev == 0) place.assign(1, beadn);
    if (S.n_prev > 1) return fail(c, RHG_ERR_UNSUPPORTED, "chain: a growth step with more than one previous bead");
    if ((int)place.size() > kChainPlace) return fail(c, RHG_ERR_UNSUPPORTED, "chain: more than 4 branches at one bead");
    S.n_place = (int)place.size();
    std::vector<char> done(n, 0), tbp(n, 0);
    for (int a : placed) done[a] = 1;
    for (int a : place) done[a] = tbp[a] = 1;
    auto code = [&](int a) {
      for (int jj = 0; jj < S.n_place; ++jj)
        if (place[jj] == a) return jj;
      if (a == cur) return kLocCur;
      if (a == S.prev) return kLocPrev;
      return kLocGlobal + a;
    };
    for (int jj = 0; jj < S.n_place; ++jj) {
      S.place[jj] = place[jj];
      S.type[jj] = ch->type[place[jj]];
      S.ref_len[jj] = ref_len(cur, place[jj]);
      S.max_bond[jj] = per_atom(ch->max_bond_change, place[jj]);
    }
    S.max_bend = per_atom(ch->max_bend_change, cur);
    S.max_cone = per_atom(ch->max_cone_rotation, cur);
    for (int jj = 0; jj < S.n_place; ++jj) {
      const int A = std::min(place[jj], cur), B = std::max(place[jj], cur);
      for (int j = 0; j < ch->n_bonds; ++j) {
        const int32_t *b = ch->bonds + 3 * j;
        if (!((A == b[0] && B == b[1]) || (A == b[1] && B == b[0]))) continue;
        if (S.n_bonds == kChainBonds) return fail(c, RHG_ERR_UNSUPPORTED, "chain: too many bonds in one step");
        const int q = S.n_bonds++;
        S.bond_a[q] = code(b[0]);
        S.bond_b[q] = code(b[1]);
        S.bond_kind[q] = b[2];
        S.bond_branch[q] = jj;
        S.bond_p0[q] = b[2] == RHG_BOND_HARMONIC ? ch->bond_par[4 * j] / E2K : ch->bond_par[4 * j];
        S.bond_p1[q] = ch->bond_par[4 * j + 1];
      }
    }
    if (S.n_prev == 0 && S.n_bonds == 0) return fail(c, RHG_ERR_INVALID, "chain: no bond to grow a bead along");
    for (int j = 0; j < ch->n_bends; ++j) {
      const int32_t *b = ch->bends + 4 * j;
      if (!(done[b[0]] && done[b[1]] && done[b[2]] && (tbp[b[0]] || tbp[b[1]] || tbp[b[2]]))) continue;
      if (S.n_bends == kChainBends) return fail(c, RHG_ERR_UNSUPPORTED, "chain: too many bends in one step");
      const int q = S.n_bends++;
      S.bend_a[q] = code(b[0]);
      S.bend_b[q] = code(b[1]);
      S.bend_c[q] = code(b[2]);
      S.bend_p0[q] = ch->bend_par[4 * j] / E2K;
      S.bend_p1[q] = ch->bend_par[4 * j + 1] * (M_PI / 180.0);
    }
    for (int j = 0; j < ch->n_torsions; ++j) {
      const int32_t *t = ch->torsions + 5 * j;
      if (!(done[t[0]] && done[t[1]] && done[t[2]] && done[t[3]] && (tbp[t[0]] || tbp[t[1]] || tbp[t[2]] || tbp[t[3]])))
        continue;
      if (S.n_tors == kChainTors) return fail(c, RHG_ERR_UNSUPPORTED, "chain: too many torsions in one step");
      const int q = S.n_tors++;
      S.tors_a[q] = code(t[0]);
      S.tors_b[q] = code(t[1]);
      S.tors_c[q] = code(t[2]);
      S.tors_d[q] = code(t[3]);
      for (int r = 0; r < 4; ++r) S.tors_p[q][r] = ch->torsion_par[4 * j + r] / E2K;
    }
    for (int j = 0; j < ch->n_intra_vdw; ++j) {
      const int A = ch->intra_vdw[2 * j], B = ch->intra_vdw[2 * j + 1];
      if (!(done[A] && done[B] && (tbp[A] || tbp[B]))) continue;
      if (S.n_intra == kChainIntra) return fail(c, RHG_ERR_UNSUPPORTED, "chain: too many intra pairs in one step");
      const int q = S.n_intra++;
      S.intra_a[q] = code(A);
      S.intra_b[q] = code(B);
      const PseudoAtom &ta = c->types[ch->type[A]], &tb = c->types[ch->type[B]];
      const double e = sqrt((ta.eps_k * c->u.K2E) * (tb.eps_k * c->u.K2E)); // src/potentials.c:3164-3171
      const double sg = 0.5 * (ta.sigma + tb.sigma);
      S.intra_scal[q] = scal[j];
      S.intra_c4[q] = 4.0 * e;
      S.intra_s2[q] = sg * sg;
      S.intra_shift[q] = c->shifted ? lj_ref(e, sg, 0.0, c->rc2_vdw) : 0.0;
    }
    for (int a : place) {
      placed.push_back(a);
      for (int ip = 0; ip < n; ++ip) todo[a][ip] = 0;
    }
    steps.push_back(S);
  }
  return RHG_OK;
}

int chain_buffers(rhg_ctx *c, const rhg_chain *ch, int64_t cap, ChainWork &w) {
  memset(&w, 0, sizeof(w));
  w.cap = cap;
  const size_t d8 = sizeof(double);
  const size_t nfb = (size_t)cap * ch->k_fb, nat = (size_t)cap * ch->n_atoms;
  const size_t ntr = (size_t)cap * ch->k_trial, ntp = ntr * kChainPlace;
#define CB(field, name, bytes)                          \
  w.field = (decltype(w.field))work(c, name, (bytes)); \
  if (!w.field) return fail(c, RHG_ERR_OOM, "chain work " name);
  CB(fbx, "c_fbx", nfb * d8);
  CB(fby, "c_fby", nfb * d8);
  CB(fbz, "c_fbz", nfb * d8);
  CB(fb_uv, "c_fbuv", nfb * d8);
  CB(W, "c_W", cap * d8);
  CB(dU, "c_dU", cap * d8);
  CB(active, "c_active", cap * 4);
  CB(n_active, "c_nactive", 64);
  CB(alive, "c_alive", cap * 4);
  CB(sW, "c_sW", cap * d8);
  CB(uhost, "c_uhost", cap * d8);
  CB(uintra, "c_uintra", cap * d8);
  CB(ubond, "c_ubond", cap * d8);
  CB(ubend, "c_ubend", cap * d8);
  CB(utors, "c_utors", cap * d8);
  CB(W1, "c_W1", cap * d8);
  CB(px, "c_px", nat * d8);
  CB(py, "c_py", nat * d8);
  CB(pz, "c_pz", nat * d8);
  CB(bx, "c_bx", (size_t)cap * kChainPlace * d8);
  CB(by, "c_by", (size_t)cap * kChainPlace * d8);
  CB(bz, "c_bz", (size_t)cap * kChainPlace * d8);
  CB(bl, "c_bl", cap * d8);
  CB(tx, "c_tx", ntp * d8);
  CB(ty, "c_ty", ntp * d8);
  CB(tz, "c_tz", ntp * d8);
  CB(tuv, "c_tuv", ntp * d8);
  CB(tRT, "c_tRT", ntr * d8);
  CB(tbond, "c_tbond", ntr * d8);
  CB(tbend, "c_tbend", ntr * d8);
  CB(ttors, "c_ttors", ntr * d8);
  CB(tintra, "c_tintra", ntr * d8);
#undef CB
  return RHG_OK;
}

// One chunk of flexible-chain insertions [first, first + n): first-bead
// trials and energies (shared with the rigid path), then per growth step
// the internal walk, the trials, their framework energies and Rosen's
// selection, then the finalisation and block sums.
int chain_chunk(rhg_ctx *c, const rhg_chain *ch, const std::vector<ChainStep> &steps, const ChainStep *d_steps,
                ChainWork &w, uint64_t seed, int64_t first, int64_t n, double *d_rec, double *d_xyz) {
  hipStream_t st = c->stream;
  w.rec = d_rec;
  w.xyz = d_xyz;
  launch_zero_counters(w.n_active, w.n_active, st);
  // first-bead trial positions: the rigid path's generator
  MolDev m;
  memset(&m, 0, sizeof(m));
  m.n_sites = 1;
  m.k_fb = ch->k_fb;
  m.k_or = 1;
  WidomArgs wa = widom_args(c, m, seed, n);
  WidomWork ww;
  memset(&ww, 0, sizeof(ww));
  ww.cap = w.cap;
  ww.fbx = w.fbx;
  ww.fby = w.fby;
  ww.fbz = w.fbz;
  {
    Timed t(c, "widom_gen");
    launch_widom_fb_gen(wa, ww, st);
  }
  const int type0 = ch->type[ch->start];
  {
    EnergyArgs ea = energy_args(c, type0, 1.0, 1, 0);
    SiteList s{w.fbx, w.fby, w.fbz, n * ch->k_fb, wa.n_eff, ch->k_fb, w.cap};
    EnergyOut o{w.fb_uv, nullptr, nullptr};
    if (int rc = site_energy(c, type0, ea, s, o, "cfb")) return rc;
  }
  ChainArgs a;
  memset(&a, 0, sizeof(a));
  a.cell = c->cell;
  a.pockets = c->pockets;
  a.beta = c->beta;
  a.seed = seed;
  a.first = first;
  a.n = n;
  a.rc2 = c->rc2_vdw;
  a.n_atoms = ch->n_atoms;
  a.start = ch->start;
  a.k_fb = ch->k_fb;
  a.k = ch->k_trial;
  a.kt = ch->k_torsion;
  a.moves = ch->moves_per_open_bead;
  {
    Timed t(c, "chain_select");
    launch_chain_fb_select(a, w, st);
  }
  for (size_t s = 0; s < steps.size(); ++s) {
    const ChainStep &S = steps[s];
    a.step = (int)s + 1;
    a.st = d_steps + s;
    {
      Timed t(c, "chain_grow");
      launch_chain_grow(a, w, S, st);
    }
    {
      Timed t(c, "chain_trials");
      launch_chain_trials(a, w, S, st);
    }
    // framework VDW of every trial, one site list per placed bead
    for (int jj = 0; jj < S.n_place; ++jj) {
      EnergyArgs ea = energy_args(c, S.type[jj], 1.0, 1, 0);
      const int64_t base = (int64_t)jj * ch->k_trial * w.cap;
      SiteList sl{w.tx + base, w.ty + base, w.tz + base, n * ch->k_trial, w.n_active, ch->k_trial, w.cap};
      EnergyOut o{w.tuv + base, nullptr, nullptr};
      if (int rc = site_energy(c, S.type[jj], ea, sl, o, "cgr")) return rc;
    }
    {
      Timed t(c, "chain_select");
      launch_chain_select(a, w, st);
    }
  }
  {
    Timed t(c, "chain_select");
    launch_chain_finalize(a, w, st);
  }
  {
    Timed t(c, "block_reduce");
    double *part = (double *)work(c, "c_part", ((w.cap + kSeg - 1) / kSeg + 1) * 10 * sizeof(double));
    if (!part) return fail(c, RHG_ERR_OOM, "chain reduce buffers");
    launch_block_reduce(wa, w.W, w.dU, part, kSeg, c->d_acc, st);
  }
  launch_advance(c->d_ctl, n, st);
  if (hipGetLastError() != hipSuccess) return fail(c, RHG_ERR_HIP, "chain launch failed");
  return RHG_OK;
}

} // namespace

extern "C" {

int rhg_widom_run(rhg_ctx *c, const rhg_molecule *mol, uint64_t seed, int64_t n_total, int64_t first,
                  int64_t n, int32_t steps_per_cycle, rhg_widom_blocks *blocks, double *rec) {
  if (int rc = check_range(c, mol, n_total, first, n, steps_per_cycle)) return rc;
  MolDev m;
  if (int rc = build_mol(c, mol, m)) return rc;
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  // sums of earlier submits that were never collected stay out of this call
  if (int rc = widom_state(c)) return rc;
  fill(c->d_acc, 15, 0.0, c->stream);
  for (double &v : c->ig_pending) v = 0.0;
  if (int rc = widom_enqueue(c, mol, m, seed, n_total, first, n, steps_per_cycle, rec)) return rc;
  return widom_collect(c, blocks);
}

int rhg_pore_size_distribution(rhg_ctx *c, double probe_factor, int64_t cycles, int64_t n_b, uint64_t seed,
                               int32_t bins, double range, uint64_t *hist, double *largest) {
  if (!c || !hist) return RHG_ERR_INVALID;
  if (cycles < 0 || n_b < 0 || bins < 1 || !(range > 0.0) || !(probe_factor > 0.0))
    return fail(c, RHG_ERR_INVALID, "pore-size distribution: bad cycles, centres, bins, range or factor");
  if (cycles == 0) return RHG_OK;
  hipSetDevice(c->device);
  hipStream_t st = c->stream;
  const int n = c->n;
  std::vector<double4> at((size_t)std::max(n, 1));
  for (int j = 0; j < n; ++j) at[j] = make_double4(c->x[j], c->y[j], c->z[j], 0.5 * probe_factor * c->sig[j]);
  double4 *d_at = (double4 *)work(c, "psd_atoms", at.size() * sizeof(double4));
  const int64_t chunk = std::min<int64_t>(cycles, 32768);
  constexpr int kPerThread = 8;
  const int b_per_block = 256 * kPerThread;
  const int blocks_per_cycle = (int)std::max<int64_t>(1, (n_b + b_per_block - 1) / b_per_block);
  double *ax = (double *)work(c, "psd_a", (size_t)chunk * 4 * sizeof(double));
  double *part = (double *)work(c, "psd_part", (size_t)chunk * blocks_per_cycle * sizeof(double));
  double *lg = (double *)work(c, "psd_largest", (size_t)chunk * sizeof(double));
  unsigned long long *cnt = (unsigned long long *)work(c, "psd_counts", (size_t)bins * sizeof(unsigned long long));
  if (!d_at || !ax || !part || !lg || !cnt) return fail(c, RHG_ERR_OOM, "pore-size distribution buffers");
  HIPC(c, hipMemcpyAsync(d_at, at.data(), at.size() * sizeof(double4), hipMemcpyHostToDevice, st));
  fill((double *)cnt, bins, 0.0, st); // all-zero bits
  PsdArgs a;
  memset(&a, 0, sizeof(a));
  a.cell = c->cell;
  a.pockets = c->pockets;
  a.atoms = d_at;
  a.n = n;
  a.seed = seed;
  a.n_b = n_b;
  a.b_per_block = b_per_block;
  a.blocks_per_cycle = blocks_per_cycle;
  a.bins = bins;
  a.delta = range / bins;
  for (int64_t f = 0; f < cycles; f += chunk) {
    a.first = f;
    a.n_cycles = std::min(chunk, cycles - f);
    {
      Timed t(c, "psd");
      launch_psd(a, ax, part, lg, cnt, st);
    }
    if (largest)
      HIPC(c, hipMemcpyAsync(largest + f, lg, (size_t)a.n_cycles * sizeof(double), hipMemcpyDeviceToHost, st));
  }
  HIPC(c, hipGetLastError());
  std::vector<unsigned long long> h((size_t)bins);
  HIPC(c, hipMemcpyAsync(h.data(), cnt, (size_t)bins * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  HIPC(c, hipStreamSynchronize(st));
  // the reference adds 1 to every bin up to a cycle's index (src/sample.c:3514-3516)
  uint64_t acc = 0;
  for (int k = bins - 1; k >= 0; --k) {
    acc += h[k];
    hist[k] += acc;
  }
  return RHG_OK;
}

int rhg_widom_chain_run(rhg_ctx *c, const rhg_chain *ch, uint64_t seed, int64_t n_total, int64_t first,
                        int64_t n, int32_t steps_per_cycle, rhg_widom_blocks *blocks, double *rec, double *xyz) {
  static_assert(kCrecN == RHG_CREC_N && kCrecStage == RHG_CREC_STAGE && kCrecW1 == RHG_CREC_W1,
                "chain record layout");
  if (!c || !ch) return RHG_ERR_INVALID;
  if (n_total < 0 || first < 0 || n < 0 || first + n > n_total || steps_per_cycle < 1)
    return fail(c, RHG_ERR_INVALID, "bad insertion range");
  std::vector<ChainStep> steps;
  if (int rc = chain_plan(c, ch, steps)) return rc;
  if (c->ads_n > 0)
    return fail(c, RHG_ERR_UNSUPPORTED, "chain: adsorbates are present (the flexible-chain path is at infinite dilution)");
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  hipStream_t st = c->stream;
  if (int rc = widom_state(c)) return rc;
  fill(c->d_acc, 15, 0.0, st);
  for (double &v : c->ig_pending) v = 0.0;
  ChainStep *d_steps = (ChainStep *)work(c, "c_steps", steps.size() * sizeof(ChainStep));
  if (!d_steps) return fail(c, RHG_ERR_OOM, "chain step tables");
  HIPC(c, hipMemcpyAsync(d_steps, steps.data(), steps.size() * sizeof(ChainStep), hipMemcpyHostToDevice, st));
  const int64_t cap = std::min<int64_t>(n, c->chunk);
  ChainWork w;
  if (int rc = chain_buffers(c, ch, cap, w)) return rc;
  double *d_rec = nullptr, *d_xyz = nullptr;
  if (rec && !(d_rec = (double *)work(c, "c_rec", (size_t)cap * kCrecN * sizeof(double))))
    return fail(c, RHG_ERR_OOM, "chain records");
  if (xyz && !(d_xyz = (double *)work(c, "c_xyz", (size_t)cap * ch->n_atoms * 3 * sizeof(double))))
    return fail(c, RHG_ERR_OOM, "chain positions");
  int64_t jb[6];
  block_bounds(n_total, steps_per_cycle, jb);
  launch_set_ctl(c->d_ctl, first, first + n, jb, st);
  for (int64_t off = 0; off < n; off += cap) {
    const int64_t cnt = std::min(cap, n - off);
    if (int rc = check_fault(c)) return rc;
    note_ideal_gas(c, jb, first + off, cnt);
    if (d_xyz) fill(d_xyz, cnt * ch->n_atoms * 3, 0.0, st);
    if (int rc = chain_chunk(c, ch, steps, d_steps, w, seed, first + off, cnt, d_rec, d_xyz)) return rc;
    if (rec)
      HIPC(c, hipMemcpyAsync(rec + off * kCrecN, d_rec, (size_t)cnt * kCrecN * sizeof(double),
                             hipMemcpyDeviceToHost, st));
    if (xyz)
      HIPC(c, hipMemcpyAsync(xyz + off * ch->n_atoms * 3, d_xyz, (size_t)cnt * ch->n_atoms * 3 * sizeof(double),
                             hipMemcpyDeviceToHost, st));
  }
  return widom_collect(c, blocks);
}

int rhg_widom_submit(rhg_ctx *c, const rhg_molecule *mol, uint64_t seed, int64_t n_total, int64_t first,
                     int64_t n, int32_t steps_per_cycle) {
  if (int rc = check_range(c, mol, n_total, first, n, steps_per_cycle)) return rc;
  MolDev m;
  if (int rc = build_mol(c, mol, m)) return rc;
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  return widom_enqueue(c, mol, m, seed, n_total, first, n, steps_per_cycle, nullptr);
}

int rhg_widom_query(rhg_ctx *c) {
  if (!c) return RHG_ERR_INVALID;
  hipSetDevice(c->device);
  hipError_t e = hipStreamQuery(c->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(c, RHG_ERR_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int rhg_widom_collect(rhg_ctx *c, rhg_widom_blocks *blocks) {
  if (!c) return RHG_ERR_INVALID;
  hipSetDevice(c->device);
  return widom_collect(c, blocks);
}

int rhg_set_widom_options(rhg_ctx *c, int64_t chunk, int32_t graphs) {
  if (!c) return RHG_ERR_INVALID;
  if (chunk < 1 || chunk > ((int64_t)1 << 24)) return fail(c, RHG_ERR_INVALID, "chunk out of range [1, 2^24]");
  c->chunk = chunk;
  c->graphs = graphs ? 1 : 0;
  return RHG_OK;
}

int rhg_inject_fault(rhg_ctx *c, int64_t after_chunks) {
  if (!c) return RHG_ERR_INVALID;
  c->fault_after = after_chunks < 0 ? -1 : after_chunks;
  return RHG_OK;
}

int rhg_widom_run_trials(rhg_ctx *c, const rhg_molecule *mol, int64_t n, const double *fb,
                         const double *cord, const double *u_sel, double *rec) {
  if (!c || !mol || n < 0 || !rec || (n > 0 && (!fb || !u_sel))) return RHG_ERR_INVALID;
  MolDev m;
  if (int rc = build_mol(c, mol, m)) return rc;
  if (m.n_sites > 1 && !cord) return fail(c, RHG_ERR_INVALID, "cord required");
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  if (int rc = widom_state(c)) return rc;
  WidomBufs B;
  if (int rc = widom_buffers(c, m, n, B)) return rc;
  size_t bfb = (size_t)n * m.k_fb * 3 * sizeof(double);
  size_t bcord = (size_t)n * m.k_or * m.n_sites * 3 * sizeof(double);
  size_t bsel = (size_t)n * 2 * sizeof(double);
  double *hfb = (double *)work(c, "h_fb", bfb), *hsel = (double *)work(c, "h_sel", bsel);
  double *hcord = m.n_sites > 1 ? (double *)work(c, "h_cord", bcord) : nullptr;
  if (!hfb || !hsel || (m.n_sites > 1 && !hcord)) return fail(c, RHG_ERR_OOM, "trial buffers");
  upload(c, hfb, fb, bfb);
  upload(c, hsel, u_sel, bsel);
  if (hcord) upload(c, hcord, cord, bcord);
  B.w.h_fb = hfb;
  B.w.h_cord = hcord;
  B.w.h_usel = hsel;
  int64_t jb[6] = {0, n, n, n, n, n};
  launch_set_ctl(c->d_ctl, 0, n, jb, c->stream);
  WidomArgs a = widom_args(c, m, 0, n);
  a.use_host_trials = 1;
  if (int rc = widom_chunk(c, mol, m, B, a, false)) return rc;
  double *drec = (double *)work(c, "w_rec", (size_t)n * RHG_REC_N * sizeof(double));
  if (!drec) return fail(c, RHG_ERR_OOM, "rec buffer");
  launch_widom_records(a, B.w, drec, c->stream);
  HIPC(c, hipMemcpyAsync(rec, drec, (size_t)n * RHG_REC_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

int rhg_widom_retrace_trials(rhg_ctx *c, const rhg_molecule *mol, const int32_t *exclude, int64_t n,
                             const double *old_xyz, const double *fb, const double *cord, double *rec) {
  if (!c || !mol || n < 0 || !rec || (n > 0 && (!old_xyz || !fb))) return RHG_ERR_INVALID;
  std::vector<int32_t> ex((size_t)std::max<int64_t>(n, 1), c->ads_nmol);
  if (exclude)
    for (int64_t i = 0; i < n; ++i) {
      if (exclude[i] >= c->ads_nmol) return fail(c, RHG_ERR_INVALID, "no such adsorbate molecule");
      ex[i] = exclude[i] < 0 ? c->ads_nmol : exclude[i];
    }
  MolDev m;
  if (int rc = build_mol(c, mol, m)) return rc;
  if (m.n_sites > 1 && !cord) return fail(c, RHG_ERR_INVALID, "cord required");
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  if (int rc = widom_state(c)) return rc;
  WidomBufs B;
  if (int rc = widom_buffers(c, m, n, B)) return rc;
  // first-bead trial 0 is the molecule's starting bead (src/cbmc.c:2990)
  std::vector<double> hfb_h(fb, fb + (size_t)n * m.k_fb * 3);
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) hfb_h[(size_t)i * m.k_fb * 3 + d] = old_xyz[((size_t)i * m.n_sites + m.start) * 3 + d];
  size_t bfb = (size_t)n * m.k_fb * 3 * sizeof(double);
  size_t bcord = (size_t)n * m.k_or * m.n_sites * 3 * sizeof(double);
  size_t bold = (size_t)n * m.n_sites * 3 * sizeof(double);
  double *hfb = (double *)work(c, "h_fb", bfb), *hold = (double *)work(c, "h_old", bold);
  double *hcord = m.n_sites > 1 ? (double *)work(c, "h_cord", bcord) : nullptr;
  int32_t *hex = (int32_t *)work(c, "h_excl", (size_t)n * sizeof(int32_t));
  if (!hfb || !hold || !hex || (m.n_sites > 1 && !hcord)) return fail(c, RHG_ERR_OOM, "trial buffers");
  upload(c, hfb, hfb_h.data(), bfb);
  upload(c, hold, old_xyz, bold);
  upload(c, hex, ex.data(), (size_t)n * sizeof(int32_t));
  if (hcord) upload(c, hcord, cord, bcord);
  B.w.h_fb = hfb;
  B.w.h_cord = hcord;
  B.w.h_usel = nullptr;
  B.w.h_old = hold;
  int64_t jb[6] = {0, n, n, n, n, n};
  launch_set_ctl(c->d_ctl, 0, n, jb, c->stream);
  WidomArgs a = widom_args(c, m, 0, n);
  a.use_host_trials = 1;
  a.retrace = 1;
  a.ewald = 0; // RetraceMolecule has no Fourier step
  a.excl = hex;
  if (int rc = widom_chunk(c, mol, m, B, a, false)) return rc;
  double *drec = (double *)work(c, "w_rec", (size_t)n * RHG_REC_N * sizeof(double));
  if (!drec) return fail(c, RHG_ERR_OOM, "rec buffer");
  launch_widom_records(a, B.w, drec, c->stream);
  HIPC(c, hipMemcpyAsync(rec, drec, (size_t)n * RHG_REC_N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_; // hfb_h and ex are released on return
  return RHG_OK;
}

int rhg_widom_estimators(const rhg_widom_blocks *b, double T, double rho, double igw, double out[16]) {
  if (!b || !out || T <= 0.0 || rho <= 0.0 || igw <= 0.0) return RHG_ERR_INVALID;
  Units u = make_units();
  const int NB = RHG_NR_BLOCKS;
  double sum = 0.0, sum2 = 0.0;
  for (int i = 0; i < NB; i++) {
    double tmp = 0.0;
    if (b->count[i] > 0.0) {
      tmp = (1.0 / (kMolarGas * T * rho)) * b->sum_w[i] / (igw * b->count[i]);
      sum += tmp;
      sum2 += tmp * tmp;
    }
    out[i] = tmp;
  }
  out[5] = sum / (double)NB;
  out[6] = 2.0 * sqrt(fabs((sum2 / (double)NB) - (sum * sum) / (double)(NB * NB)));
  sum = sum2 = 0.0;
  for (int i = 0; i < NB; i++) {
    double tmp = 0.0;
    if (b->count[i] > 0.0) {
      tmp = (b->sum_wdu[i] / b->sum_w[i] - 0.0 / b->count[i]) * u.E2K;
      sum += tmp;
      sum2 += tmp * tmp;
    }
    out[7 + i] = tmp;
  }
  out[12] = sum / (double)NB;
  out[13] = 2.0 * sqrt(fabs((sum2 / (double)NB) - (sum * sum) / (double)(NB * NB)));
  double cnt = 0.0, sw = 0.0;
  for (int i = 0; i < NB; i++)
    if (b->count[i] > 0.0) {
      sw += b->sum_w[i] / igw;
      cnt += b->count[i];
    }
  out[14] = cnt > 0.0 ? (1.0 / (kMolarGas * T * rho)) * (sw / cnt) : 0.0;
  out[15] = cnt > 0.0 ? sw * igw / cnt : 0.0;
  return RHG_OK;
}

int rhg_ewald_fourier_adsorbate(rhg_ctx *c, int64_t n_mol, int32_t n_sites, const int32_t *types,
                                const double *xyz, double *d_host, double *d_ads) {
  if (!c || n_mol < 0 || n_sites < 1 || n_sites > RHG_MAX_SITES || !types || (n_mol > 0 && !xyz))
    return RHG_ERR_INVALID;
  if (c->charge_method != RHG_CHARGE_EWALD) {
    for (int64_t i = 0; i < n_mol; ++i) {
      if (d_host) d_host[i] = 0.0;
      if (d_ads) d_ads[i] = 0.0;
    }
    return RHG_OK;
  }
  for (int j = 0; j < n_sites; ++j)
    if (int rc = check_type(c, types[j])) return rc;
  if (n_mol == 0) return RHG_OK;
  hipSetDevice(c->device);
  std::vector<double> px((size_t)n_sites * n_mol), py(px.size()), pz(px.size());
  for (int64_t i = 0; i < n_mol; ++i)
    for (int j = 0; j < n_sites; ++j) {
      px[j * n_mol + i] = xyz[(i * n_sites + j) * 3 + 0];
      py[j * n_mol + i] = xyz[(i * n_sites + j) * 3 + 1];
      pz[j * n_mol + i] = xyz[(i * n_sites + j) * 3 + 2];
    }
  size_t b = px.size() * sizeof(double);
  double *dx = (double *)work(c, "f_x", b), *dy = (double *)work(c, "f_y", b), *dz = (double *)work(c, "f_z", b);
  double *dh = (double *)work(c, "f_h", n_mol * sizeof(double)), *da = (double *)work(c, "f_a", n_mol * sizeof(double));
  if (!dx || !dy || !dz || !dh || !da) return fail(c, RHG_ERR_OOM, "fourier buffers");
  upload(c, dx, px.data(), b);
  upload(c, dy, py.data(), b);
  upload(c, dz, pz.data(), b);
  if (int rc = ensure_fourier(c)) return rc;
  FourierArgs f;
  memset(&f, 0, sizeof(f));
  f.e = ewald_dev(c);
  f.n_sites = n_sites;
  for (int j = 0; j < n_sites; ++j) {
    double q = c->types[types[j]].q;
    if (considered_charged(q)) {
      f.cidx[f.nc] = j;
      f.q[f.nc] = 1.0 * q;
      f.self_energy += c->u.C * (f.q[f.nc] * f.q[f.nc]) * c->alpha / sqrt(M_PI);
      f.nc++;
    }
  }
  f.px = dx;
  f.py = dy;
  f.pz = dz;
  f.stride = n_mol;
  f.n = n_mol;
  f.n_dev = nullptr;
  f.d_host = dh;
  f.d_ads = da;
  f.sads = c->ads_n > 0 ? c->d_sads : nullptr;
  f.net_q_ads = c->net_q_ads;
  if (f.nc == 0) {
    fill(dh, n_mol, 0.0, c->stream);
    fill(da, n_mol, 0.0, c->stream);
  } else {
    double2 *part = (double2 *)work(c, "f_part", (size_t)(c->kx + 1) * n_mol * sizeof(double2));
    if (!part) return fail(c, RHG_ERR_OOM, "fourier partials");
    Timed t(c, "fourier");
    launch_fourier(f, part, c->stream);
  }
  HIPC(c, hipGetLastError());
  if (d_host) HIPC(c, hipMemcpyAsync(d_host, dh, n_mol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (d_ads) HIPC(c, hipMemcpyAsync(d_ads, da, n_mol * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

// --------------------------------------------------------- finite loading
int rhg_set_adsorbates(rhg_ctx *c, int32_t n_mol, const int32_t *start, const int32_t *types,
                       const double *xyz) {
  if (!c || n_mol < 0 || (n_mol > 0 && (!start || !types || !xyz))) return RHG_ERR_INVALID;
  int n = 0;
  if (n_mol > 0) {
    if (start[0] != 0) return fail(c, RHG_ERR_INVALID, "molecule 0 must start at atom 0");
    for (int m = 0; m < n_mol; ++m)
      if (start[m + 1] < start[m]) return fail(c, RHG_ERR_INVALID, "molecule starts must not decrease");
    n = start[n_mol];
    for (int i = 0; i < n; ++i)
      if (int rc = check_type(c, types[i])) return rc;
  }
  hipSetDevice(c->device);
  c->ads_nmol = n_mol;
  c->ads_n = n;
  c->ads_start.assign(start, start + (n_mol > 0 ? n_mol + 1 : 0));
  c->ads_type.assign(types, types + n);
  c->ads_xyz.assign(xyz, xyz + 3 * (size_t)n);
  c->net_q_ads = 0.0;
  if (n > 0) {
    std::vector<int32_t> mol((size_t)n);
    for (int m = 0; m < n_mol; ++m)
      for (int a = start[m]; a < start[m + 1]; ++a) mol[a] = m;
    if (c->ads_mol_cap < (size_t)n) {
      dfree(c, c->d_ads_mol);
      c->d_ads_mol = nullptr;
      c->ads_mol_cap = 0;
      if (int rc = dalloc(c, (void **)&c->d_ads_mol, (size_t)n * sizeof(int32_t))) return rc;
      c->ads_mol_cap = (size_t)n;
    }
    if (int rc = upload(c, c->d_ads_mol, mol.data(), (size_t)n * sizeof(int32_t))) return rc;
    if (int rs_ = sync_stream(c)) return rs_;
    for (auto &t : c->types)
      if (int rc = ads_records(c, t)) return rc;
    // StoreTotalChargeAdsorbates, NetChargeAdsorbates
    std::vector<double> fx, fy, fz, qq;
    for (int a = 0; a < n; ++a) {
      const double q = c->types[types[a]].q;
      if (!considered_charged(q)) continue;
      double f[3];
      frac_2pi(c, xyz + 3 * a, f);
      fx.push_back(f[0]);
      fy.push_back(f[1]);
      fz.push_back(f[2]);
      qq.push_back(q);
      c->net_q_ads += q;
    }
    if (c->charge_method == RHG_CHARGE_EWALD) {
      if (!c->d_sads)
        if (int rc = dalloc(c, (void **)&c->d_sads, (size_t)std::max(c->nk, 1) * sizeof(double2))) return rc;
      if (int rc = structure_factor(c, fx, fy, fz, qq, c->d_sads)) return rc;
    }
  }
  c->gen++; // captured Widom chunks hold the adsorbate arguments
  return RHG_OK;
}

int rhg_inter_energy_at_positions(rhg_ctx *c, int64_t n, const double *xyz, int32_t type, int32_t exclude,
                                  double *u_vdw, double *u_cc) {
  if (!c || n < 0 || (n > 0 && !xyz)) return RHG_ERR_INVALID;
  if (int rc = check_type(c, type)) return rc;
  if (n == 0) return RHG_OK;
  if (c->ads_n == 0) {
    for (int64_t i = 0; i < n; ++i) {
      if (u_vdw) u_vdw[i] = 0.0;
      if (u_cc) u_cc[i] = 0.0;
    }
    return RHG_OK;
  }
  hipSetDevice(c->device);
  std::vector<double> sx((size_t)n), sy((size_t)n), sz((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    sx[i] = xyz[3 * i];
    sy[i] = xyz[3 * i + 1];
    sz[i] = xyz[3 * i + 2];
  }
  const size_t b = (size_t)n * sizeof(double);
  double *dx = (double *)work(c, "e_x", b), *dy = (double *)work(c, "e_y", b), *dz = (double *)work(c, "e_z", b);
  double *dv = (double *)work(c, "e_v", b), *dc = (double *)work(c, "e_c", b);
  if (!dx || !dy || !dz || !dv || !dc) return fail(c, RHG_ERR_OOM, "work buffers");
  upload(c, dx, sx.data(), b);
  upload(c, dy, sy.data(), b);
  upload(c, dz, sz.data(), b);
  const AdsArgs a = ads_args(c, type, exclude, 1, 1);
  if (!a.do_cc) fill(dc, n, 0.0, c->stream);
  SiteList sl{dx, dy, dz, n, nullptr, 1, 0};
  launch_site_ads(a, sl, dv, a.do_cc ? dc : nullptr, c->stream);
  HIPC(c, hipGetLastError());
  if (u_vdw) HIPC(c, hipMemcpyAsync(u_vdw, dv, b, hipMemcpyDeviceToHost, c->stream));
  if (u_cc) HIPC(c, hipMemcpyAsync(u_cc, dc, b, hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

int rhg_ewald_fourier_move(rhg_ctx *c, int32_t n_new, const int32_t *new_types, const double *new_xyz,
                           int32_t old_mol, double out[2]) {
  if (!c || !out || n_new < 0 || (n_new > 0 && (!new_types || !new_xyz))) return RHG_ERR_INVALID;
  if (old_mol >= c->ads_nmol) return fail(c, RHG_ERR_INVALID, "no such adsorbate molecule");
  out[0] = out[1] = 0.0;
  if (c->charge_method != RHG_CHARGE_EWALD) return RHG_OK;
  for (int j = 0; j < n_new; ++j)
    if (int rc = check_type(c, new_types[j])) return rc;
  // charged sites: the new molecule's, then the removed molecule's with
  // negated charges, so that one structure-factor sum is S_new - S_old
  std::vector<double> px, py, pz, q;
  std::vector<int> grp;
  double self_new = 0.0, self_old = 0.0;
  for (int j = 0; j < n_new; ++j) {
    const double qj = c->types[new_types[j]].q;
    if (!considered_charged(qj)) continue;
    px.push_back(new_xyz[3 * j]);
    py.push_back(new_xyz[3 * j + 1]);
    pz.push_back(new_xyz[3 * j + 2]);
    q.push_back(1.0 * qj);
    grp.push_back(0);
    self_new += c->u.C * (qj * qj) * c->alpha / sqrt(M_PI); // src/ewald.c:6973
  }
  if (old_mol >= 0)
    for (int a = c->ads_start[old_mol]; a < c->ads_start[old_mol + 1]; ++a) {
      const double qa = c->types[c->ads_type[a]].q;
      if (!considered_charged(qa)) continue;
      px.push_back(c->ads_xyz[3 * a]);
      py.push_back(c->ads_xyz[3 * a + 1]);
      pz.push_back(c->ads_xyz[3 * a + 2]);
      q.push_back(-qa);
      grp.push_back(1);
      self_old += c->u.C * (qa * qa) * c->alpha / sqrt(M_PI); // src/ewald.c:6950
    }
  const int nc = (int)q.size();
  if (nc > kMaxSites) return fail(c, RHG_ERR_UNSUPPORTED, "more than 8 charged sites in one move");
  if (nc == 0) return RHG_OK;
  hipSetDevice(c->device);
  const size_t b = (size_t)nc * sizeof(double);
  double *dx = (double *)work(c, "m_x", b), *dy = (double *)work(c, "m_y", b), *dz = (double *)work(c, "m_z", b);
  double *dr = (double *)work(c, "m_out", 2 * sizeof(double));
  double2 *part = (double2 *)work(c, "m_part", (size_t)(c->kx + 1) * sizeof(double2));
  if (!dx || !dy || !dz || !dr || !part) return fail(c, RHG_ERR_OOM, "fourier move buffers");
  upload(c, dx, px.data(), b);
  upload(c, dy, py.data(), b);
  upload(c, dz, pz.data(), b);
  if (int rc = ensure_fourier(c)) return rc;
  FourierArgs f;
  memset(&f, 0, sizeof(f));
  f.e = ewald_dev(c);
  f.n_sites = nc;
  f.nc = nc;
  for (int t = 0; t < nc; ++t) {
    f.cidx[t] = t;
    f.q[t] = q[t];
    f.grp[t] = grp[t];
  }
  f.self_energy = self_new - self_old;
  f.px = dx;
  f.py = dy;
  f.pz = dz;
  f.stride = 1;
  f.n = 1;
  f.n_dev = nullptr;
  f.d_host = dr;
  f.d_ads = dr + 1;
  f.sads = c->ads_n > 0 ? c->d_sads : nullptr;
  f.net_q_ads = c->net_q_ads;
  {
    Timed t(c, "fourier");
    launch_fourier(f, part, c->stream);
  }
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(out, dr, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

// ------------------------------------------------------------ surface area
int rhg_surface_area(rhg_ctx *c, int32_t probe_type, double distance_factor, int64_t samples, uint64_t seed,
                     const uint8_t *interacts, double *area, uint64_t *counts) {
  if (!c || !area || samples < 1 || !(distance_factor > 0.0)) return RHG_ERR_INVALID;
  if (int rc = check_type(c, probe_type)) return rc;
  *area = 0.0;
  if (c->n == 0) return RHG_OK;
  if ((double)c->n * (double)samples > 9.0e18) return fail(c, RHG_ERR_INVALID, "too many sample points");
  hipSetDevice(c->device);
  const PseudoAtom &t = c->types[probe_type];
  // equilibrium distance per atom: factor x the LB-mixed sigma of the pair
  // (PotentialParms[typeA][typeB][1], src/mc_moves.c:16458)
  std::vector<double4> h((size_t)c->n);
  std::vector<double> d((size_t)c->n);
  // an atom whose pseudo-atom does not interact gets distance 0: no sphere
  // of its own, and rr < 0 never excludes a point
  for (int i = 0; i < c->n; ++i) {
    d[i] = (!interacts || interacts[i]) ? distance_factor * (0.5 * (t.sigma + c->sig[i])) : 0.0;
    h[i] = make_double4(c->x[i], c->y[i], c->z[i], d[i]);
  }
  double4 *da = (double4 *)work(c, "sa_atoms", (size_t)c->n * sizeof(double4));
  unsigned long long *dc = (unsigned long long *)work(c, "sa_counts", (size_t)c->n * sizeof(unsigned long long));
  if (!da || !dc) return fail(c, RHG_ERR_OOM, "surface area buffers");
  upload(c, da, h.data(), h.size() * sizeof(double4));
  fill((double *)dc, c->n, 0.0, c->stream); // all-zero bits
  SurfaceArgs a;
  memset(&a, 0, sizeof(a));
  a.cell = c->cell;
  a.pockets = c->pockets;
  a.atoms = da;
  a.n = c->n;
  a.samples = samples;
  a.seed = seed;
  {
    Timed tm(c, "surface_area");
    launch_surface_area(a, dc, c->stream);
  }
  HIPC(c, hipGetLastError());
  std::vector<unsigned long long> hc((size_t)c->n);
  HIPC(c, hipMemcpyAsync(hc.data(), dc, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  // sum_i (counted/total) 4 pi d_i^2 in atom order (src/mc_moves.c:16478-16480)
  double sum = 0.0;
  for (int i = 0; i < c->n; ++i) {
    const bool on = !interacts || interacts[i];
    if (on) sum += ((double)hc[i] / (double)samples) * 4.0 * M_PI * (d[i] * d[i]);
    if (counts) counts[i] = on ? hc[i] : 0;
  }
  *area = sum;
  return RHG_OK;
}

// ------------------------------------------------------------------ grids
namespace {

// Grid geometry of MakeGrid (src/grids.c:460-498, 587-598): the box that
// encloses the unit cell, N = (int)(size / spacing) made odd, Delta = size / N
void grid_geometry(rhg_grid *g, const rhg_ctx *c, const double unit_box[9], double spacing) {
  double size[3] = {0.0, 0.0, 0.0}, shift[3] = {0.0, 0.0, 0.0};
  const int col[3][3] = {{AX, AY, AZ}, {BX, BY, BZ}, {CX, CY, CZ}};
  for (int v = 0; v < 3; v++)
    for (int d = 0; d < 3; d++) {
      size[d] += fabs(unit_box[col[v][d]]);
      if (unit_box[col[v][d]] < 0.0) shift[d] += unit_box[col[v][d]];
    }
  for (int d = 0; d < 3; d++) {
    g->n[d] = (int)(size[d] / spacing);
    if (g->n[d] % 2 == 0) g->n[d]++;
    g->size[d] = size[d];
    g->shift[d] = shift[d];
    g->delta[d] = size[d] / (double)g->n[d];
  }
  g->spacing = spacing;
  memcpy(g->ubox, unit_box, 9 * sizeof(double));
  invert3x3(g->ubox, g->uinv);
  double up[7];
  cell_properties(g->ubox, up);
  for (int d = 0; d < 3; d++) {
    g->ulen[d] = up[d];
    g->n_cells[d] = std::max(1, (int)lround(c->props[d] / up[d]));
  }
  g->bc = c->cell.bc;
}

int64_t grid_nodes(const rhg_grid *g) { return (int64_t)(g->n[0] + 1) * (g->n[1] + 1) * (g->n[2] + 1); }

GridArgs grid_args(const rhg_ctx *c, const rhg_grid *g) {
  GridArgs a;
  memset(&a, 0, sizeof(a));
  a.cell = c->cell;
  for (int d = 0; d < 3; ++d) {
    a.n[d] = g->n[d];
    a.size[d] = g->size[d];
    a.shift[d] = g->shift[d];
    a.delta[d] = g->delta[d];
  }
  a.n_atoms = c->n;
  a.fp64 = g->fp64;
  a.out = g->d;
  return a;
}

// framework atoms as {x, y, z, q} for the real-space Coulomb kernels
double4 *atoms_xyzq(rhg_ctx *c) {
  double4 *d = (double4 *)work(c, "aq", (size_t)std::max(c->n, 1) * sizeof(double4));
  if (!d) return nullptr;
  std::vector<double4> h((size_t)std::max(c->n, 1));
  for (int i = 0; i < c->n; ++i) h[i] = make_double4(c->x[i], c->y[i], c->z[i], c->q[i]);
  upload(c, d, h.data(), h.size() * sizeof(double4));
  return d;
}

// Every node of the grid, in launches of at most 2^22 nodes.  With cell
// lists on, the nodes of a launch are binned into the cell grid like trial
// sites and each tile walks its cell's lists; precision 1 (VDW grids only)
// uses the image lists.
int build_nodes(rhg_ctx *c, rhg_grid *g, GridArgs &a, const float4 *par) {
  const int64_t nodes = grid_nodes(g);
  const int64_t step = (int64_t)1 << 22;
  const bool culled = c->cull && c->n > 0;
  if (culled) {
    if (int rc = ensure_cell_grid(c)) return rc;
    if (!a.coulomb && c->precision != RHG_PREC_FP64 && c->use_images)
      if (int rc = ensure_image_list(c)) return rc;
  }
  const int prec = (!a.coulomb && c->precision != RHG_PREC_FP64 && c->use_images && c->cg.img_off) ? 1 : 0;
  unsigned long long *pairs =
      culled ? (unsigned long long *)work(c, "grid_pairs", 2 * kPairLanes * sizeof(unsigned long long)) : nullptr;
  if (culled && !pairs) return fail(c, RHG_ERR_OOM, "grid pair counter");
  if (pairs) fill((double *)pairs, 2 * kPairLanes, 0.0, c->stream); // all-zero bits
  for (int64_t n0 = 0; n0 < nodes; n0 += step) {
    a.node0 = n0;
    a.n_nodes = std::min(step, nodes - n0);
    if (!culled) {
      Timed tm(c, "grid_build");
      launch_grid_build(a, c->stream);
      continue;
    }
    const size_t nb = (size_t)step * sizeof(double);
    double *gx = (double *)work(c, "g_x", nb), *gy = (double *)work(c, "g_y", nb),
           *gz = (double *)work(c, "g_z", nb);
    if (!gx || !gy || !gz) return fail(c, RHG_ERR_OOM, "grid node buffers");
    launch_grid_nodes(a, gx, gy, gz, c->stream);
    SiteList sl{gx, gy, gz, a.n_nodes, nullptr, 1, 0};
    BinWork b;
    const int ncell = c->cg.ncell;
    b.cell_of = (int32_t *)work(c, "g_cellof", (size_t)step * sizeof(int32_t));
    b.perm = (int32_t *)work(c, "g_perm", (size_t)step * sizeof(int32_t));
    // cleared in full when allocated; each binning's scan clears what it used
    b.count = (int32_t *)work(c, "g_count", (size_t)ncell * sizeof(int32_t), true);
    b.counts_zero = 1;
    b.rank = (int32_t *)work(c, "g_rank", (size_t)step * sizeof(int32_t));
    b.start = (int32_t *)work(c, "g_start", (size_t)(ncell + 1) * sizeof(int32_t));
    b.tile_off = (int32_t *)work(c, "g_tiles", (size_t)(ncell + 1) * sizeof(int32_t));
    b.pairs = pairs;
    b.fix = nullptr;
    if (!b.cell_of || !b.perm || !b.count || !b.rank || !b.start || !b.tile_off)
      return fail(c, RHG_ERR_OOM, "grid binning scratch");
    launch_bin_sites(c->cell, c->cg, sl, b, c->stream);
    Timed tm(c, "grid_build");
    launch_grid_culled(a, c->cg, sl, b, par, prec, c->stream);
  }
  HIPC(c, hipGetLastError());
  std::vector<unsigned long long> cnt(2 * kPairLanes, 0);
  if (pairs)
    HIPC(c, hipMemcpyAsync(cnt.data(), pairs, cnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  unsigned long long total = 0;
  for (unsigned long long v : cnt) total += v;
  g->exec_pairs = pairs ? (double)total : (double)nodes * (double)c->n;
  return RHG_OK;
}

int grid_alloc(rhg_ctx *c, rhg_grid *g) {
  g->bytes = (size_t)grid_nodes(g) * 8 * (g->fp64 ? sizeof(double) : sizeof(float));
  hipSetDevice(c->device);
  if (hipMalloc(&g->d, g->bytes) != hipSuccess) {
    g->d = nullptr;
    return fail(c, RHG_ERR_OOM, "grid allocation");
  }
  return RHG_OK;
}

} // namespace

int rhg_make_grid(rhg_ctx *c, int32_t type, const double unit_box[9], double spacing, int32_t fp64,
                  rhg_grid **out) {
  if (!c || !unit_box || !out || !(spacing > 0.0)) return RHG_ERR_INVALID;
  if (int rc = check_type(c, type)) return rc;
  *out = nullptr;
  // src/grids.c:441-449
  if (std::min(c->props[3], std::min(c->props[4], c->props[5])) < c->rc_vdw)
    return fail(c, RHG_ERR_INVALID, "cutoff larger than a perpendicular box width");
  std::unique_ptr<rhg_grid> g(new rhg_grid());
  g->ctx = c;
  g->device = c->device;
  grid_geometry(g.get(), c, unit_box, spacing);
  g->fp64 = fp64 ? 1 : 0;
  if (int rc = grid_alloc(c, g.get())) return rc;
  const PseudoAtom &t = c->types[type];
  GridArgs a = grid_args(c, g.get());
  a.rc2 = c->rc2_vdw;
  a.a64 = t.d64;
  a.c24 = t.c24;
  a.c96 = t.c96;
  a.c384 = t.c384;
  if (int rc = build_nodes(c, g.get(), a, t.par)) {
    rhg_grid_free(g.release());
    return rc;
  }
  *out = g.release();
  return RHG_OK;
}

int rhg_make_coulomb_grid(rhg_ctx *c, const double unit_box[9], double spacing, int32_t fp64,
                          rhg_grid **out) {
  if (!c || !unit_box || !out || !(spacing > 0.0)) return RHG_ERR_INVALID;
  *out = nullptr;
  // src/grids.c:600-614: Ewald only here (Wolf's method is not a charge
  // method of this library)
  if (c->charge_method != RHG_CHARGE_EWALD)
    return fail(c, RHG_ERR_UNSUPPORTED, "the Coulomb grid needs the Ewald charge method");
  if (std::min(c->props[3], std::min(c->props[4], c->props[5])) < c->rc_cc)
    return fail(c, RHG_ERR_INVALID, "cutoff larger than a perpendicular box width");
  std::unique_ptr<rhg_grid> g(new rhg_grid());
  g->ctx = c;
  g->device = c->device;
  grid_geometry(g.get(), c, unit_box, spacing);
  g->fp64 = fp64 ? 1 : 0;
  g->coulomb = 1;
  g->ewald_precision = c->ewald_precision;
  if (int rc = grid_alloc(c, g.get())) return rc;
  GridArgs a = grid_args(c, g.get());
  a.rc2 = c->rc2_cc;
  a.coulomb = 1;
  a.cc = c->u.C;
  a.alpha = c->alpha;
  a.sqrtpi = sqrt(M_PI);
  a.aq = atoms_xyzq(c);
  if (!a.aq) {
    rhg_grid_free(g.release());
    return fail(c, RHG_ERR_OOM, "atom buffer");
  }
  if (int rc = build_nodes(c, g.get(), a, nullptr)) {
    rhg_grid_free(g.release());
    return rc;
  }
  *out = g.release();
  return RHG_OK;
}

int rhg_grid_info(const rhg_grid *g, double out[13]) {
  if (!g || !out) return RHG_ERR_INVALID;
  for (int d = 0; d < 3; ++d) {
    out[d] = g->n[d];
    out[3 + d] = g->size[d];
    out[6 + d] = g->shift[d];
    out[9 + d] = g->delta[d];
  }
  out[12] = 8.0 * (g->fp64 ? 8.0 : 4.0);
  return RHG_OK;
}

int rhg_grid_download(const rhg_grid *g, void *host, size_t bytes) {
  if (!g || !host || bytes < g->bytes) return RHG_ERR_INVALID;
  hipSetDevice(g->ctx->device);
  if (hipMemcpy(host, g->d, g->bytes, hipMemcpyDeviceToHost) != hipSuccess) return RHG_ERR_HIP;
  return RHG_OK;
}

namespace {
int interpolate(rhg_grid *g, int64_t n, const double *xyz, double charge, double *u) {
  rhg_ctx *c = g->ctx;
  hipSetDevice(c->device);
  double *dx = (double *)work(c, "i_xyz", (size_t)n * 3 * sizeof(double));
  double *du = (double *)work(c, "i_u", (size_t)n * sizeof(double));
  if (!dx || !du) return fail(c, RHG_ERR_OOM, "interp buffers");
  upload(c, dx, xyz, (size_t)n * 3 * sizeof(double));
  const InterpArgs a = interp_args(g, charge);
  {
    Timed t(c, "grid_interp");
    launch_interp(a, dx, n, du, c->stream);
  }
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(u, du, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}
} // namespace

int rhg_interpolate_vdw_grid(rhg_grid *g, int64_t n, const double *xyz, double *u) {
  if (!g || n < 0 || (n > 0 && (!xyz || !u))) return RHG_ERR_INVALID;
  if (g->coulomb) return fail(g->ctx, RHG_ERR_INVALID, "a Coulomb grid is interpolated with its charge");
  if (n == 0) return RHG_OK;
  return interpolate(g, n, xyz, 1.0, u);
}

int rhg_interpolate_coulomb_grid(rhg_grid *g, int64_t n, const double *xyz, double charge, double *u) {
  if (!g || n < 0 || (n > 0 && (!xyz || !u))) return RHG_ERR_INVALID;
  if (!g->coulomb) return fail(g->ctx, RHG_ERR_INVALID, "not a Coulomb grid");
  if (n == 0) return RHG_OK;
  return interpolate(g, n, xyz, charge, u);
}

// ------------------------------------------------------------ grid files
// WriteVDWGrid / WriteCoulombGrid (src/grids.c:692-742, 815-876): the
// header fields one after the other (spacing, N, SizeGrid, ShiftGrid,
// Delta, UnitCellSize, NumberOfUnitCells, and for a Coulomb grid the Ewald
// precision), then the float channels channel-major, m, i, j, k.
namespace {
constexpr size_t kVdwHeader = 8 + 12 + 4 * 24 + 12;
constexpr size_t kCoulombHeader = kVdwHeader + 8;
} // namespace

int rhg_grid_write(const rhg_grid *g, const char *path, const int32_t n_unit_cells[3]) {
  if (!g || !path) return RHG_ERR_INVALID;
  rhg_ctx *c = g->ctx;
  if (g->fp64) return fail(c, RHG_ERR_INVALID, "grid files hold FP32 channels; build the grid with fp64 = 0");
  const int64_t nodes = grid_nodes(g);
  hipSetDevice(c->device);
  float *tmp = nullptr;
  if (hipMalloc(&tmp, g->bytes) != hipSuccess) return fail(c, RHG_ERR_OOM, "grid transpose buffer");
  std::vector<float> host((size_t)nodes * 8);
  launch_grid_transpose((const float *)g->d, tmp, nodes, 1, c->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host.data(), tmp, g->bytes, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(tmp);
  if (e != hipSuccess) return fail(c, RHG_ERR_HIP, std::string("grid transpose: ") + hipGetErrorString(e));
  FILE *f = fopen(path, "wb");
  if (!f) return fail(c, RHG_ERR_IO, std::string("cannot open ") + path + " for writing");
  const int32_t nc[3] = {n_unit_cells ? n_unit_cells[0] : g->n_cells[0],
                         n_unit_cells ? n_unit_cells[1] : g->n_cells[1],
                         n_unit_cells ? n_unit_cells[2] : g->n_cells[2]};
  const int32_t N[3] = {g->n[0], g->n[1], g->n[2]};
  bool ok = fwrite(&g->spacing, sizeof(double), 1, f) == 1 && fwrite(N, sizeof(int32_t), 3, f) == 3 &&
            fwrite(g->size, sizeof(double), 3, f) == 3 && fwrite(g->shift, sizeof(double), 3, f) == 3 &&
            fwrite(g->delta, sizeof(double), 3, f) == 3 && fwrite(g->ulen, sizeof(double), 3, f) == 3 &&
            fwrite(nc, sizeof(int32_t), 3, f) == 3;
  if (ok && g->coulomb) ok = fwrite(&g->ewald_precision, sizeof(double), 1, f) == 1;
  if (ok) ok = fwrite(host.data(), sizeof(float), host.size(), f) == host.size();
  if (fclose(f) != 0) ok = false;
  if (!ok) return fail(c, RHG_ERR_IO, std::string("short write to ") + path);
  return RHG_OK;
}

int rhg_grid_read(rhg_ctx *c, const char *path, int32_t coulomb, const double unit_box[9],
                  const int32_t n_unit_cells[3], rhg_grid **out) {
  if (!c || !path || !unit_box || !out) return RHG_ERR_INVALID;
  *out = nullptr;
  FILE *f = fopen(path, "rb");
  if (!f) return fail(c, RHG_ERR_IO, std::string("file ") + path + " does not exist");
  std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
  std::unique_ptr<rhg_grid> g(new rhg_grid());
  g->ctx = c;
  g->device = c->device;
  int32_t N[3], nc[3];
  double ucs[3], ep = -1.0;
  bool ok = fread(&g->spacing, sizeof(double), 1, f) == 1 && fread(N, sizeof(int32_t), 3, f) == 3 &&
            fread(g->size, sizeof(double), 3, f) == 3 && fread(g->shift, sizeof(double), 3, f) == 3 &&
            fread(g->delta, sizeof(double), 3, f) == 3 && fread(ucs, sizeof(double), 3, f) == 3 &&
            fread(nc, sizeof(int32_t), 3, f) == 3;
  if (ok && coulomb) ok = fread(&ep, sizeof(double), 1, f) == 1;
  if (!ok) return fail(c, RHG_ERR_IO, std::string("short grid header in ") + path);
  for (int d = 0; d < 3; ++d)
    if (N[d] < 1 || N[d] > (1 << 20) || !(g->size[d] > 0.0))
      return fail(c, RHG_ERR_IO, std::string("malformed grid header in ") + path);
  for (int d = 0; d < 3; ++d) g->n[d] = N[d];
  const int64_t nodes = grid_nodes(g.get());
  const size_t header = coulomb ? kCoulombHeader : kVdwHeader;
  if (fseek(f, 0, SEEK_END) != 0) return fail(c, RHG_ERR_IO, "seek");
  const long len = ftell(f);
  if (len < 0 || (size_t)len != header + (size_t)nodes * 8 * sizeof(float))
    return fail(c, RHG_ERR_IO, std::string("grid file size does not match its header: ") + path);
  fseek(f, (long)header, SEEK_SET);
  // src/grids.c:787-793, 931-947
  double up[7];
  cell_properties(unit_box, up);
  for (int d = 0; d < 3; ++d)
    if (fabs(up[d] - ucs[d]) > 1e-8)
      return fail(c, RHG_ERR_INVALID, "error in reading the grid: unit-cell size does not match");
  if (coulomb && n_unit_cells)
    for (int d = 0; d < 3; ++d)
      if (nc[d] != n_unit_cells[d])
        return fail(c, RHG_ERR_INVALID, "mismatch of the number of unit cells");
  memcpy(g->ubox, unit_box, 9 * sizeof(double));
  invert3x3(g->ubox, g->uinv);
  for (int d = 0; d < 3; ++d) {
    g->ulen[d] = up[d];
    g->n_cells[d] = nc[d];
  }
  g->bc = c->cell.bc;
  g->fp64 = 0;
  g->coulomb = coulomb ? 1 : 0;
  g->ewald_precision = ep;
  std::vector<float> host((size_t)nodes * 8);
  if (fread(host.data(), sizeof(float), host.size(), f) != host.size())
    return fail(c, RHG_ERR_IO, std::string("short grid data in ") + path);
  if (int rc = grid_alloc(c, g.get())) return rc;
  float *tmp = nullptr;
  if (hipMalloc(&tmp, g->bytes) != hipSuccess) {
    rhg_grid_free(g.release());
    return fail(c, RHG_ERR_OOM, "grid transpose buffer");
  }
  hipError_t e = hipMemcpyAsync(tmp, host.data(), g->bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch_grid_transpose(tmp, (float *)g->d, nodes, 0, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(tmp);
  if (e != hipSuccess) {
    rhg_grid_free(g.release());
    return fail(c, RHG_ERR_HIP, std::string("grid upload: ") + hipGetErrorString(e));
  }
  *out = g.release();
  return RHG_OK;
}

int rhg_use_grid(rhg_ctx *c, int32_t type, rhg_grid *g) {
  if (!c) return RHG_ERR_INVALID;
  if (int rc = check_type(c, type)) return rc;
  if (g && (g->coulomb || g->device != c->device))
    return fail(c, RHG_ERR_INVALID, "a VDW grid on the context's device is needed");
  if ((int)c->vdw_grids.size() < (int)c->types.size()) c->vdw_grids.resize(c->types.size(), nullptr);
  c->vdw_grids[type] = g;
  c->gen++; // captured chunks hold the energy kernels
  return RHG_OK;
}

int rhg_use_coulomb_grid(rhg_ctx *c, rhg_grid *g) {
  if (!c) return RHG_ERR_INVALID;
  if (g && (!g->coulomb || g->device != c->device))
    return fail(c, RHG_ERR_INVALID, "a Coulomb grid on the context's device is needed");
  if (g && c->charge_method != RHG_CHARGE_EWALD)
    return fail(c, RHG_ERR_UNSUPPORTED, "the Coulomb grid needs the Ewald charge method");
  c->coulomb_grid = g;
  c->gen++;
  return RHG_OK;
}

int rhg_grid_meta(const rhg_grid *g, double out[10]) {
  if (!g || !out) return RHG_ERR_INVALID;
  out[0] = g->spacing;
  for (int d = 0; d < 3; ++d) {
    out[1 + d] = g->ulen[d];
    out[4 + d] = g->n_cells[d];
  }
  out[7] = g->coulomb;
  out[8] = g->ewald_precision;
  out[9] = g->exec_pairs;
  return RHG_OK;
}

void rhg_grid_free(rhg_grid *g) {
  if (!g) return;
  if (g->d) {
    hipSetDevice(g->device);
    hipFree(g->d);
  }
  delete g;
}

int rhg_derivatives_at_positions_vdw(rhg_ctx *c, int32_t type, int64_t n, const double *xyz, double *out) {
  if (!c || n < 0 || (n > 0 && (!xyz || !out))) return RHG_ERR_INVALID;
  if (int rc = check_type(c, type)) return rc;
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  double *dx = (double *)work(c, "d_xyz", (size_t)n * 3 * sizeof(double));
  double *dout = (double *)work(c, "d_out", (size_t)n * 14 * sizeof(double));
  if (!dx || !dout) return fail(c, RHG_ERR_OOM, "derivative buffers");
  upload(c, dx, xyz, (size_t)n * 3 * sizeof(double));
  const PseudoAtom &t = c->types[type];
  GridArgs a;
  memset(&a, 0, sizeof(a));
  a.cell = c->cell;
  a.rc2 = c->rc2_vdw;
  a.n_atoms = c->n;
  a.a64 = t.d64;
  a.c24 = t.c24;
  a.c96 = t.c96;
  a.c384 = t.c384;
  launch_derivatives(a, dx, n, dout, c->stream);
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(out, dout, (size_t)n * 14 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

int rhg_derivatives_at_positions_real(rhg_ctx *c, int64_t n, const double *xyz, double *out,
                                      double *smallest_r) {
  if (!c || n < 0 || (n > 0 && (!xyz || !out || !smallest_r))) return RHG_ERR_INVALID;
  if (c->charge_method != RHG_CHARGE_EWALD)
    return fail(c, RHG_ERR_UNSUPPORTED, "real-space derivatives need the Ewald charge method");
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  double *dx = (double *)work(c, "d_xyz", (size_t)n * 3 * sizeof(double));
  double *dout = (double *)work(c, "d_out", (size_t)n * 14 * sizeof(double));
  double *dr = (double *)work(c, "d_rmin", (size_t)n * sizeof(double));
  if (!dx || !dout || !dr) return fail(c, RHG_ERR_OOM, "derivative buffers");
  upload(c, dx, xyz, (size_t)n * 3 * sizeof(double));
  GridArgs a;
  memset(&a, 0, sizeof(a));
  a.cell = c->cell;
  a.rc2 = c->rc2_cc;
  a.n_atoms = c->n;
  a.coulomb = 1;
  a.cc = c->u.C;
  a.alpha = c->alpha;
  a.sqrtpi = sqrt(M_PI);
  a.aq = atoms_xyzq(c);
  if (!a.aq) return fail(c, RHG_ERR_OOM, "atom buffer");
  launch_derivatives_real(a, dx, n, dout, dr, c->stream);
  HIPC(c, hipGetLastError());
  HIPC(c, hipMemcpyAsync(out, dout, (size_t)n * 14 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipMemcpyAsync(smallest_r, dr, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

// -------------------------------------------------------------- profiling
int rhg_profile_enable(rhg_ctx *c, int32_t on) {
  if (!c) return RHG_ERR_INVALID;
  if (on < 0 || on > 2) return fail(c, RHG_ERR_INVALID, "profiling mode must be 0, 1 or 2");
  c->prof = on;
  return RHG_OK;
}

int rhg_profile_get(rhg_ctx *c, const char *kernel, int64_t *launches, double *total_ms) {
  if (!c || !kernel) return RHG_ERR_INVALID;
  hipSetDevice(c->device);
  drain_profile(c);
  auto it = c->stats.find(kernel);
  if (launches) *launches = it == c->stats.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.second;
  return RHG_OK;
}

int rhg_profile_reset(rhg_ctx *c) {
  if (!c) return RHG_ERR_INVALID;
  drain_profile(c);
  c->stats.clear();
  return RHG_OK;
}

int rhg_widom_stats(rhg_ctx *c, double out[12], int32_t reset) {
  if (!c || !out) return RHG_ERR_INVALID;
  hipSetDevice(c->device);
  double st[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<unsigned long long> lanes(3 * kPairLanes, 0);
  DevBuf &pb = c->bufs["pair_counter"];
  if (c->d_acc) HIPC(c, hipMemcpyAsync(st, c->d_acc + 16, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  if (pb.p)
    HIPC(c, hipMemcpyAsync(lanes.data(), pb.p, lanes.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
  if (reset) {
    if (c->d_acc) fill(c->d_acc + 16, 9, 0.0, c->stream);
    if (pb.p) fill((double *)pb.p, 3 * kPairLanes, 0.0, c->stream); // all-zero bits
  }
  if (int rs_ = sync_stream(c)) return rs_;
  for (int i = 0; i < 9; ++i) out[i] = st[i];
  for (int k = 0; k < 3; ++k) {
    unsigned long long t = 0;
    for (int l = 0; l < kPairLanes; ++l) t += lanes[k * kPairLanes + l];
    out[9 + k] = (double)t;
  }
  return RHG_OK;
}

int rhg_set_tiles(rhg_ctx *c, int32_t sites_per_lane, int32_t waves_per_block) {
  if (!c) return RHG_ERR_INVALID;
  if (sites_per_lane < 1 || sites_per_lane > kImgLane)
    return fail(c, RHG_ERR_INVALID, "sites per lane out of range [1, 4]");
  if (waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4)
    return fail(c, RHG_ERR_INVALID, "waves per block must be 1, 2 or 4");
  c->tile_sites = 64 * sites_per_lane;
  c->waves = waves_per_block;
  c->gen++; // captured chunks hold the launch shapes
  return RHG_OK;
}

int rhg_set_culling(rhg_ctx *c, int32_t on, double cell_size) {
  if (!c) return RHG_ERR_INVALID;
  if (on && !(cell_size > 0.0)) return fail(c, RHG_ERR_INVALID, "cell size must be > 0");
  hipSetDevice(c->device);
  if (on < 0 || on > 2) return fail(c, RHG_ERR_INVALID, "culling mode must be 0, 1 or 2");
  c->cull = on ? 1 : 0;
  c->use_images = on == 2 ? 1 : 0;
  c->gen++;
  if (on && cell_size != c->cell_size) {
    c->cell_size = cell_size;
    c->grid_ready = false;
  }
  return RHG_OK;
}

int rhg_set_block_pockets(rhg_ctx *c, int32_t n, const double *xyz, const double *radius,
                          const double unit_box[9], int32_t invert) {
  if (!c || n < 0 || (n > 0 && (!xyz || !radius || !unit_box))) return RHG_ERR_INVALID;
  hipSetDevice(c->device);
  if (int rs_ = sync_stream(c)) return rs_;
  dfree(c, c->d_pockets);
  c->d_pockets = nullptr;
  memset(&c->pockets, 0, sizeof(c->pockets));
  c->gen++; // captured chunks hold the pocket arguments
  if (n == 0) return RHG_OK;
  std::vector<double> h((size_t)n * 4);
  for (int i = 0; i < n; ++i) {
    h[4 * i] = xyz[3 * i];
    h[4 * i + 1] = xyz[3 * i + 1];
    h[4 * i + 2] = xyz[3 * i + 2];
    h[4 * i + 3] = radius[i];
  }
  if (int rc = dalloc(c, (void **)&c->d_pockets, h.size() * sizeof(double))) return rc;
  HIPC(c, hipMemcpy(c->d_pockets, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  c->pockets.n = n;
  c->pockets.invert = invert ? 1 : 0;
  c->pockets.bc = c->cell.bc == RHG_BC_RECTANGULAR ? 1 : 2;
  c->pockets.pk = c->d_pockets;
  memcpy(c->pockets.ubox, unit_box, 9 * sizeof(double));
  invert3x3(c->pockets.ubox, c->pockets.uinv);
  return RHG_OK;
}

int rhg_blocked_pocket(rhg_ctx *c, int64_t n, const double *xyz, uint8_t *blocked) {
  if (!c || n < 0 || (n > 0 && (!xyz || !blocked))) return RHG_ERR_INVALID;
  if (n == 0) return RHG_OK;
  hipSetDevice(c->device);
  const size_t b = (size_t)n * 3 * sizeof(double);
  double *dx = (double *)work(c, "p_xyz", b);
  uint8_t *df = (uint8_t *)work(c, "p_flags", (size_t)n);
  if (!dx || !df) return fail(c, RHG_ERR_OOM, "pocket buffers");
  upload(c, dx, xyz, b);
  launch_blocked_pockets(c->pockets, dx, n, df, c->stream);
  HIPC(c, hipMemcpyAsync(blocked, df, (size_t)n, hipMemcpyDeviceToHost, c->stream));
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

int rhg_workspace_create(int device, rhg_workspace **out) {
  if (!out) return RHG_ERR_INVALID;
  static std::atomic<uint64_t> next_uid{1};
  *out = new rhg_workspace();
  (*out)->device = device;
  (*out)->uid = next_uid.fetch_add(1);
  return RHG_OK;
}

void rhg_workspace_destroy(rhg_workspace *ws) {
  if (!ws) return;
  hipSetDevice(ws->device);
  if (ws->owner) {
    if (ws->owner->stream) hipStreamSynchronize(ws->owner->stream);
    ws->owner->ws = nullptr;
  }
  for (auto &b : ws->bufs) {
    hipFree(b.second.p);
    for (auto &r : b.second.retired) hipFree(r.first);
  }
  delete ws;
}

int rhg_workspace_footprint(const rhg_workspace *ws, size_t *bytes) {
  if (!ws || !bytes) return RHG_ERR_INVALID;
  size_t b = 0;
  for (auto &kv : ws->bufs) b += kv.second.bytes;
  *bytes = b;
  return RHG_OK;
}

int rhg_attach_workspace(rhg_ctx *c, rhg_workspace *ws) {
  if (!c) return RHG_ERR_INVALID;
  if (ws && ws->device != c->device) return fail(c, RHG_ERR_INVALID, "workspace on another device");
  hipSetDevice(c->device);
  if (c->ws == ws) {
    if (ws) ws->owner = c;
    return RHG_OK;
  }
  if (c->ws) {
    // the old workspace may still be in use by this context's stream
    if (c->stream) {
      if (int rs_ = sync_stream(c)) return rs_;
    }
    if (c->ws->owner == c) c->ws->owner = nullptr;
    c->ws = nullptr;
  }
  if (ws) {
    if (ws->owner && ws->owner != c) {
      if (ws->owner->stream) HIPC(c, hipStreamSynchronize(ws->owner->stream));
      ws->owner->ws = nullptr;
    }
    ws->owner = c;
    c->ws = ws;
  }
  return RHG_OK;
}

int rhg_stream_copy(int device, int64_t bytes, int32_t reps) {
  if (bytes < 16 || bytes % 16 || reps < 1) return RHG_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return RHG_ERR_HIP;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, (size_t)bytes) != hipSuccess) return RHG_ERR_OOM;
  if (hipMalloc(&b, (size_t)bytes) != hipSuccess) {
    hipFree(a);
    return RHG_ERR_OOM;
  }
  int rc = RHG_OK;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = RHG_ERR_HIP;
  for (int r = 0; r < reps && rc == RHG_OK; ++r) launch_stream_copy(a, b, bytes, st);
  if (rc == RHG_OK && hipStreamSynchronize(st) != hipSuccess) rc = RHG_ERR_HIP;
  if (st) hipStreamDestroy(st);
  hipFree(a);
  hipFree(b);
  return rc;
}

int rhg_synchronize(rhg_ctx *c) {
  if (!c) return RHG_ERR_INVALID;
  hipSetDevice(c->device);
  if (int rs_ = sync_stream(c)) return rs_;
  return RHG_OK;
}

} // extern "C"
