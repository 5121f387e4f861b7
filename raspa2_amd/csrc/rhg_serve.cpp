// Screening service: many Widom requests over a catalogue of frameworks on
// one GPU.  The reference answers one request per process (python/raspa2.py:
// 97-110 runs `simulate` once per call; src/monte_carlo.c:626-748 runs the
// systems of a process one after another).  Here a request loop keeps a
// resident cache of framework contexts (records, k-tables, cell lists) and of
// their energy grids, admits requests by device-memory footprint, and keeps
// up to `slots` requests in flight, each on its slot's stream, refilling a
// slot as soon as its request finishes (continuous batching).  Requests may
// be submitted while the loop runs.  A request is enqueued whole: its chunks
// after the first replay one captured HIP graph, its block sums accumulate on
// the device, and the slot waits once, at the end; the result does not
// depend on what runs beside it, and a request that fails is discarded and
// run again.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/rhg.h"
#include "rhg_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct FrameworkSpec {
  std::vector<double> x, y, z, q, eps, sig, mass;
  rhg_cell cell;
  rhg_forcefield ff;
  // resident state
  rhg_ctx *ctx = nullptr;
  size_t bytes = 0;          // device bytes held: the context and its grids
  uint64_t last_used = 0;    // scheduler tick of the last use (LRU)
  bool busy = false;         // a request is running on it
  std::map<std::tuple<double, double, double>, int32_t> types;
  // tuned options of this framework's size bucket
  double cell_size = 2.5;
  int64_t chunk = (int64_t)1 << 20;
  int32_t lanes = 4, waves = 2;
  // tabulated energy grids (rhg_server_use_grids): spacing > 0 serves every
  // request on this framework from grids built on first use
  double grid_spacing = 0.0;
  std::map<std::pair<double, double>, rhg_grid *> vdw_grids; // by pseudo-atom (eps, sigma)
  rhg_grid *coulomb_grid = nullptr;
  size_t grid_bytes = 0;
};

struct Job {
  int slot = -1;             // slot running the request
  rhg_request req;
  rhg_result res;
  double arrival = 0.0;      // seconds after the loop's start
  bool started = false, finished = false;
  rhg_molecule mol;
  int64_t fault_after = -1;  // host fault hook for this request
};

struct Tuning {
  int32_t max_atoms;
  double cell_size;
  int64_t chunk;
  int32_t lanes, waves;
};

} // namespace

struct rhg_server {
  int device = 0;
  size_t budget = 0;
  int slots = 4;
  int precision = RHG_PREC_MIXED;
  int64_t slice = (int64_t)1 << 18;
  std::map<int64_t, FrameworkSpec> fw;
  std::deque<Job> jobs; // requests by id; references stay valid as requests are added
  std::vector<Tuning> tuning;
  std::string err;
  uint64_t tick = 0;
  size_t resident = 0, peak_resident = 0;
  double n_created = 0, n_evicted = 0, n_retries = 0, busy_wait_s = 0;
  double setup_s = 0, submit_s = 0, collect_s = 0; // host time in set-up, submits, collects
  double grids_built = 0, grids_evicted = 0, n_finished = 0;
  // work of the finished requests: insertions and executed pair evaluations
  // of the culled site kernels by kind (LJ, LJ + Coulomb, Coulomb)
  double work_ins = 0, work_pairs[3] = {0, 0, 0};
  // device memory of the contexts' tables (one hipMalloc per slab)
  rhg_pool *pool = nullptr;
  // per slot: the scratch buffers and the stream of the request running there
  std::vector<rhg_workspace *> ws;
  std::vector<size_t> ws_bytes;
  std::vector<char> ws_busy;
  std::vector<hipStream_t> streams;
  // request loop
  std::mutex mu;
  std::condition_variable cv;      // new requests, finished requests, stop
  std::condition_variable done_cv; // finished requests
  std::vector<int64_t> waiting;    // queued requests, by arrival
  bool running = false, stopping = false;
  Clock::time_point t0;
  std::vector<std::thread> workers;
  int n_active = 0;
};

namespace {

int sfail(rhg_server *s, int code, const std::string &m) {
  s->err = m;
  return code;
}

double secs(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// Device bytes a context of this framework holds after a Widom run of a
// probe with n_sites sites (n_charged of them charged, n_types distinct
// pseudo-atoms) in chunks of `chunk` insertions: the buffers the context
// allocates, counted as it allocates them.
// *scratch receives the part that lives in a workspace when one is attached.
size_t estimate_bytes(const rhg_framework *f, const rhg_cell *cell, const rhg_forcefield *ff,
                      int32_t n_types, int64_t chunk, int32_t n_sites, int32_t n_charged, int32_t k_fb,
                      int32_t k_or, double cell_size, int32_t precision, size_t *scratch = nullptr) {
  const double *B = cell->box;
  // volume and perpendicular widths of the cell
  double a[3] = {B[0], B[1], B[2]}, b[3] = {B[3], B[4], B[5]}, c[3] = {B[6], B[7], B[8]};
  auto cross = [](const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  double bc[3], ca[3], ab[3];
  cross(b, c, bc);
  cross(c, a, ca);
  cross(a, b, ab);
  double V = std::fabs(a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2]);
  auto norm = [](const double *u) { return std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]); };
  double w[3] = {V / norm(bc), V / norm(ca), V / norm(ab)};
  const double n = std::max(f->n_atoms, 1);
  const bool ewald = ff->charge_method == RHG_CHARGE_EWALD;
  int n_q = 0;
  for (int i = 0; i < f->n_atoms; ++i)
    if (f->charge && std::fabs(f->charge[i]) > 1e-10) ++n_q;
  // atom records per pseudo-atom (64 + 40 + 3 x 8 + 16 B) and positions (24 B)
  double bytes = n * (144.0 * std::max(n_types, 1) + 24.0);
  // Ewald: the automatic alpha / kmax rule and the spherical k cut-off
  // (src/ewald.c:450-530), counted exactly
  int kx = 0;
  if (ewald) {
    double eps = std::min(std::fabs(ff->ewald_precision > 0 ? ff->ewald_precision : 1e-6), 0.5);
    double rc = ff->cutoff_charge > 0 ? ff->cutoff_charge : ff->cutoff_vdw;
    double tol = std::sqrt(std::fabs(std::log(eps * rc)));
    double alpha = std::sqrt(std::fabs(std::log(eps * rc * tol))) / rc;
    double tol1 = std::sqrt(-std::log(eps * rc * ((2.0 * tol * alpha) * (2.0 * tol * alpha))));
    // the rule takes the lengths of the cell vectors (src/ewald.c:219-221)
    const double len[3] = {norm(a), norm(b), norm(c)};
    int k[3];
    for (int d = 0; d < 3; ++d) {
      double v = 0.25 + len[d] * alpha * tol1 / M_PI;
      k[d] = (int)(v >= 0 ? v + 0.5 : v - 0.5);
    }
    int m = std::max(k[0], std::max(k[1], k[2]));
    double cut = (1.05 * m) * (1.05 * m);
    double nk = 0;
    for (int i = 0; i <= k[0]; ++i)
      for (int j = -k[1]; j <= k[1]; ++j)
        for (int l = -k[2]; l <= k[2]; ++l) {
          double q2 = (double)(i * i + j * j + l * l);
          if (q2 != 0 && q2 < cut) nk += 1;
        }
    kx = k[0];
    // k factors, structure factors, k indices, row tables, charged-atom
    // scratch, the Fourier kernel's k records and their order
    bytes += nk * (8 + 16 + 12) + (k[0] + 1) * (2 * k[1] + 1) * 8.0 + n_q * 32.0 + (nk + 8) * (32 + 4);
  }
  // cell lists: atoms within (rc + half cell diagonal) of each cell
  double rc = std::max(ff->cutoff_vdw, ff->cutoff_charge);
  double ncell = 1;
  for (int d = 0; d < 3; ++d) ncell *= std::max(1.0, std::min(64.0, std::floor(w[d] / cell_size)));
  double reach = rc + 0.5 * std::sqrt(3.0) * cell_size;
  double per_cell = std::min(n, n / V * 4.0 / 3.0 * M_PI * reach * reach * reach);
  bytes += ncell * (per_cell * 4.0 + 8.0);
  // mixed precision adds the image lists (16 B per entry)
  if (precision == RHG_PREC_MIXED) bytes += ncell * (per_cell * 16.0 + 8.0);
  // Widom buffers per insertion: first-bead trials (5 x 8 B each), per
  // insertion scalars (132 B), orientation trials (4 x 8 B each), positions of
  // the selected sites (24 B each), placed-bead Coulomb (8 B each), Fourier
  // results (16 B) and, for a charged probe with Ewald, the partial sums of
  // each kx slice (16 B each)
  const int n_place = std::max(n_sites - 1, 1);
  const int k_or_eff = n_sites > 1 ? k_or : 1;
  double per_ins = k_fb * 40.0 + 132.0 + (double)k_or_eff * n_place * 32.0 + n_sites * 24.0 +
                   n_place * 8.0 + 16.0;
  if (ewald && n_charged > 0) per_ins += (kx + 1) * 16.0;
  // binning scratch: 8 B per binned site (40 B with the unsorted and sorted
  // FP32 offsets of mixed precision) and 16 B per cell for each list
  const double per_site = precision == RHG_PREC_MIXED ? 40.0 : 8.0;
  per_ins += k_fb * per_site;
  double lists = 1;
  if (n_sites > 1) {
    per_ins += (double)k_or * (n_sites - 1) * per_site;
    lists += 1;
    if (ewald && n_charged > 0) {
      per_ins += (double)n_charged * per_site;
      lists += 1;
    }
  }
  double ws = per_ins * (double)chunk + lists * ncell * 16.0;
  // reduction partials and small control buffers (at least 256 B each)
  ws += ((double)(chunk + 4095) / 4096.0 + 1.0) * 80.0 + 64 * 256.0;
  // the context's tables are carved from 2 MiB blocks: half a block of slack
  bytes += 1.0 * (1 << 20);
  if (scratch) *scratch = (size_t)ws;
  return (size_t)(bytes + ws);
}

// FP32 bytes of one energy grid of this framework at `spacing` (MakeGrid's
// node rule over the box as unit cell, src/grids.c:441-470: per axis the
// enclosing size over the spacing, made odd, plus one node)
size_t grid_bytes_of(const rhg_cell &cell, double spacing) {
  const double *B = cell.box;
  double nodes = 1.0;
  for (int d = 0; d < 3; ++d) {
    const double size = std::fabs(B[d]) + std::fabs(B[3 + d]) + std::fabs(B[6 + d]);
    int n = (int)(size / spacing);
    if (n % 2 == 0) n++;
    nodes *= (double)(n + 1);
  }
  return (size_t)(nodes * 8.0 * sizeof(float));
}

// grids a request of probe p still needs on framework f: VDW grids of its
// distinct pseudo-atoms and, for a charged probe with Ewald, the Coulomb grid
int grids_missing(const FrameworkSpec &f, const rhg_probe &p) {
  if (!(f.grid_spacing > 0.0)) return 0;
  std::vector<std::pair<double, double>> seen;
  int miss = 0;
  bool charged = false;
  for (int j = 0; j < p.n_sites; ++j) {
    const std::pair<double, double> key(p.eps_k[j], p.sigma[j]);
    charged = charged || std::fabs(p.charge[j]) > 1e-10;
    if (f.vdw_grids.count(key) || std::find(seen.begin(), seen.end(), key) != seen.end()) continue;
    seen.push_back(key);
    ++miss;
  }
  if (charged && f.ff.charge_method == RHG_CHARGE_EWALD && !f.coulomb_grid) ++miss;
  return miss;
}

void tuned_options(rhg_server *s, FrameworkSpec &f) {
  f.cell_size = 2.5;
  f.chunk = s->slice;
  f.lanes = 4;
  f.waves = 2;
  for (const Tuning &t : s->tuning)
    if ((int)f.x.size() <= t.max_atoms) {
      f.cell_size = t.cell_size;
      f.chunk = t.chunk;
      f.lanes = t.lanes;
      f.waves = t.waves;
      break;
    }
}

void free_grids(rhg_server *s, FrameworkSpec &f) {
  for (auto &kv : f.vdw_grids) {
    rhg_grid_free(kv.second);
    s->grids_evicted += 1;
  }
  f.vdw_grids.clear();
  if (f.coulomb_grid) {
    rhg_grid_free(f.coulomb_grid);
    s->grids_evicted += 1;
  }
  f.coulomb_grid = nullptr;
  f.grid_bytes = 0;
}

// Frees the grids of an idle framework that probe `p` does not use (VDW grids
// of other pseudo-atoms, the Coulomb grid when no site is charged), detaching
// them from the context first; returns the bytes freed.  Caller holds the lock.
size_t drop_unused_grids(rhg_server *s, FrameworkSpec &f, const rhg_probe &p) {
  const size_t gb = grid_bytes_of(f.cell, f.grid_spacing);
  size_t freed = 0;
  bool charged = false;
  for (int j = 0; j < p.n_sites; ++j) charged = charged || std::fabs(p.charge[j]) > 1e-10;
  for (auto it = f.vdw_grids.begin(); it != f.vdw_grids.end();) {
    bool used = false;
    for (int j = 0; j < p.n_sites; ++j) used = used || it->first == std::make_pair(p.eps_k[j], p.sigma[j]);
    if (used) {
      ++it;
      continue;
    }
    if (f.ctx)
      for (const auto &t : f.types)
        if (std::get<0>(t.first) == it->first.first && std::get<1>(t.first) == it->first.second)
          rhg_use_grid(f.ctx, t.second, nullptr);
    rhg_grid_free(it->second);
    it = f.vdw_grids.erase(it);
    freed += gb;
  }
  if (f.coulomb_grid && !charged) {
    if (f.ctx) rhg_use_coulomb_grid(f.ctx, nullptr);
    rhg_grid_free(f.coulomb_grid);
    f.coulomb_grid = nullptr;
    freed += gb;
  }
  s->grids_evicted += (double)(freed / std::max(gb, (size_t)1));
  f.grid_bytes -= std::min(f.grid_bytes, freed);
  f.bytes -= std::min(f.bytes, freed);
  s->resident -= std::min(s->resident, freed);
  return freed;
}

// caller holds the lock; the framework is idle (its context has no stream)
void evict(rhg_server *s, FrameworkSpec &f) {
  rhg_destroy(f.ctx); // grids are borrowed by the context: it goes first
  f.ctx = nullptr;
  free_grids(s, f);
  s->resident -= f.bytes;
  f.bytes = 0;
  f.types.clear();
  s->n_evicted += 1;
}

size_t ws_total(const rhg_server *s) {
  size_t t = 0;
  for (size_t b : s->ws_bytes) t += b;
  return t;
}

void note_peak(rhg_server *s) { s->peak_resident = std::max(s->peak_resident, s->resident + ws_total(s)); }

// Reserves device memory for a request on slot `slot` (caller holds the
// lock): the framework context's own bytes (if it is not resident yet), the
// grids it still lacks, plus the slot's scratch must fit the budget beside
// the other contexts and workspaces; idle frameworks (contexts and grids) are
// evicted least recently used first, then the framework's own grids that this
// probe does not use, then idle workspaces are released.
// *wait is set when the request would fit once running requests finish.
int reserve(rhg_server *s, FrameworkSpec &f, const rhg_probe &p, int slot, bool *wait) {
  *wait = false;
  f.last_used = ++s->tick;
  if (!f.ctx) tuned_options(s, f);
  rhg_framework fr{(int32_t)f.x.size(), f.x.data(), f.y.data(), f.z.data(), f.q.data(),
                   f.eps.data(), f.sig.data(), f.mass.data()};
  int n_charged = 0;
  for (int j = 0; j < p.n_sites; ++j) n_charged += std::fabs(p.charge[j]) > 1e-10 ? 1 : 0;
  size_t scratch = 0;
  const size_t total = estimate_bytes(&fr, &f.cell, &f.ff, p.n_sites, std::min(s->slice, f.chunk),
                                      p.n_sites, n_charged, p.k_fb, p.k_or, f.cell_size, s->precision,
                                      &scratch);
  size_t own = f.ctx ? 0 : total - scratch;
  const int miss = grids_missing(f, p);
  if (miss > 0) {
    own += (size_t)miss * grid_bytes_of(f.cell, f.grid_spacing);
    // a grid build bins up to 2^22 nodes at a time (32 B each) in the workspace
    scratch = std::max(scratch, ((size_t)1 << 22) * 32 + ((size_t)64 * 64 * 64) * 16);
  }
  // f.bytes: what the framework holds now; grids of other probes go first
  if (f.bytes + own + scratch > s->budget) drop_unused_grids(s, f, p);
  if (f.bytes + own + scratch > s->budget)
    return sfail(s, RHG_ERR_NOT_ADMITTED, "request larger than the device budget");
  auto need = [&]() {
    size_t w = ws_total(s) - s->ws_bytes[slot] + std::max(s->ws_bytes[slot], scratch);
    return s->resident + own + w;
  };
  while (need() > s->budget) {
    FrameworkSpec *lru = nullptr;
    for (auto &kv : s->fw)
      if (kv.second.ctx && !kv.second.busy && &kv.second != &f &&
          (!lru || kv.second.last_used < lru->last_used))
        lru = &kv.second;
    if (lru) {
      evict(s, *lru);
      continue;
    }
    if (drop_unused_grids(s, f, p) > 0) continue;
    // release the buffers of idle workspaces
    bool freed = false;
    for (size_t w = 0; w < s->ws.size(); ++w)
      if (!s->ws_busy[w] && s->ws_bytes[w] > 0) {
        rhg_workspace_destroy(s->ws[w]);
        rhg_workspace_create(s->device, &s->ws[w]);
        s->ws_bytes[w] = 0;
        freed = true;
      }
    if (freed) continue;
    *wait = true;
    return RHG_ERR_NOT_ADMITTED;
  }
  // reserved now; the worker builds what it stands for outside the lock, and
  // the accounting switches to measured bytes when the request ends
  f.bytes += own;
  s->resident += own;
  s->ws_bytes[slot] = std::max(s->ws_bytes[slot], scratch);
  note_peak(s);
  return RHG_OK;
}

int create_context(rhg_server *s, FrameworkSpec &f, int slot) {
  rhg_framework fr{(int32_t)f.x.size(), f.x.data(), f.y.data(), f.z.data(), f.q.data(),
                   f.eps.data(), f.sig.data(), f.mass.data()};
  int rc = rhg_create_on_stream(&fr, &f.cell, &f.ff, s->device, s->streams[slot], s->ws[slot], s->pool, &f.ctx);
  if (rc) {
    f.ctx = nullptr;
    return rc;
  }
  rhg_set_precision(f.ctx, s->precision);
  rhg_set_culling(f.ctx, 2, f.cell_size);
  rhg_set_widom_options(f.ctx, std::min(f.chunk, s->slice), 1);
  rhg_set_tiles(f.ctx, f.lanes, f.waves);
  return RHG_OK;
}

int probe_molecule(FrameworkSpec &f, const rhg_probe &p, rhg_molecule &m) {
  memset(&m, 0, sizeof(m));
  m.n_sites = p.n_sites;
  for (int j = 0; j < p.n_sites; ++j) {
    auto key = std::make_tuple(p.eps_k[j], p.sigma[j], p.charge[j]);
    auto it = f.types.find(key);
    int32_t t;
    if (it == f.types.end()) {
      if (int rc = rhg_add_pseudo_atom(f.ctx, p.eps_k[j], p.sigma[j], p.charge[j], &t)) return rc;
      f.types[key] = t;
    } else {
      t = it->second;
    }
    m.type[j] = t;
    for (int d = 0; d < 3; ++d) m.body[j][d] = p.body[j][d];
  }
  m.start = p.start;
  m.k_fb = p.k_fb;
  m.k_or = p.k_or;
  return RHG_OK;
}

// Grids of a grid-served framework for the probe's pseudo-atoms: built on
// first use over the box (FP32 nodes, the server's precision), then attached
// to every pseudo-atom of the context they serve (UseTabularGrid,
// src/framework_energy.c:73-76, 197-200).  *built receives the number built.
int ensure_grids(rhg_server *s, FrameworkSpec &f, const rhg_probe &p, const rhg_molecule &m, int *built) {
  *built = 0;
  if (!(f.grid_spacing > 0.0)) return RHG_OK;
  bool charged = false;
  for (int j = 0; j < p.n_sites; ++j) {
    const std::pair<double, double> key(p.eps_k[j], p.sigma[j]);
    charged = charged || std::fabs(p.charge[j]) > 1e-10;
    auto it = f.vdw_grids.find(key);
    if (it == f.vdw_grids.end()) {
      rhg_grid *g = nullptr;
      if (int rc = rhg_make_grid(f.ctx, m.type[j], f.cell.box, f.grid_spacing, 0, &g)) return rc;
      it = f.vdw_grids.emplace(key, g).first;
      f.grid_bytes += grid_bytes_of(f.cell, f.grid_spacing);
      ++*built;
    }
    if (int rc = rhg_use_grid(f.ctx, m.type[j], it->second)) return rc;
  }
  if (charged && f.ff.charge_method == RHG_CHARGE_EWALD) {
    if (!f.coulomb_grid) {
      if (int rc = rhg_make_coulomb_grid(f.ctx, f.cell.box, f.grid_spacing, 0, &f.coulomb_grid)) {
        f.coulomb_grid = nullptr;
        return rc;
      }
      f.grid_bytes += grid_bytes_of(f.cell, f.grid_spacing);
      ++*built;
    }
    if (int rc = rhg_use_coulomb_grid(f.ctx, f.coulomb_grid)) return rc;
  }
  (void)s;
  return RHG_OK;
}

// caller holds the lock
void finish(rhg_server *s, Job &j, FrameworkSpec &f, int rc) {
  if (rc != RHG_OK) j.res.status = rc;
  j.finished = true;
  j.res.finished_s = secs(s->t0, Clock::now());
  j.res.latency_s = j.res.finished_s - j.arrival;
  f.busy = false;
  f.last_used = ++s->tick;
  if (j.slot >= 0) {
    size_t b = 0;
    if (rhg_workspace_footprint(s->ws[j.slot], &b) == RHG_OK) s->ws_bytes[j.slot] = b;
    s->ws_busy[j.slot] = 0;
    j.slot = -1;
  }
  if (f.ctx && rc == RHG_ERR_HIP) {
    evict(s, f); // a device error may have left the context unusable
  } else if (f.ctx) {
    // from now on account for what the context and its grids really hold
    size_t b = 0;
    if (rhg_footprint(f.ctx, &b) == RHG_OK) {
      const size_t now = b + f.grid_bytes;
      s->resident = s->resident - f.bytes + now;
      f.bytes = now;
    }
  } else {
    s->resident -= f.bytes; // reserved for a context that was never created
    f.bytes = 0;
    free_grids(s, f);
  }
  note_peak(s);
  s->n_finished += 1;
  s->cv.notify_all();
  s->done_cv.notify_all();
}

// Runs one request on `slot` (no lock held): makes the framework resident on
// the slot's stream and workspace, enqueues every insertion of the request,
// and waits once for the block sums.  A failed submit discards what was
// enqueued and runs the request again (at most 3 times).
int execute(rhg_server *s, Job &j, FrameworkSpec &f, int slot, double *t_setup, double *t_sub, double *t_col,
            bool *created, int *grids, double work[12]) {
  const Clock::time_point ts = Clock::now();
  int rc = RHG_OK;
  *created = !f.ctx;
  if (*created) {
    rc = create_context(s, f, slot);
  } else {
    rc = rhg_use_stream(f.ctx, s->streams[slot]);
    if (rc == RHG_OK) rc = rhg_attach_workspace(f.ctx, s->ws[slot]);
  }
  if (rc == RHG_OK) rc = probe_molecule(f, j.req.probe, j.mol);
  if (rc == RHG_OK) rc = ensure_grids(s, f, j.req.probe, j.mol, grids);
  *t_setup = secs(ts, Clock::now());
  *t_sub = *t_col = 0.0;
  while (rc == RHG_OK) {
    if (j.fault_after >= 0) {
      rhg_inject_fault(f.ctx, j.fault_after);
      j.fault_after = -1;
    }
    const Clock::time_point tq = Clock::now();
    int sub = rhg_widom_submit(f.ctx, &j.mol, j.req.seed, j.req.n_insertions, 0, j.req.n_insertions,
                               j.req.steps_per_cycle);
    const Clock::time_point tc = Clock::now();
    *t_sub += secs(tq, tc);
    if (sub == RHG_OK) {
      rc = rhg_widom_collect(f.ctx, &j.res.blocks);
      *t_col += secs(tc, Clock::now());
      if (rc == RHG_OK) rc = rhg_widom_stats(f.ctx, work, 1);
      break;
    }
    rhg_inject_fault(f.ctx, -1);
    if (rhg_widom_collect(f.ctx, nullptr) != RHG_OK || ++j.res.retries > 3) {
      rc = sub;
      break;
    }
  }
  if (f.ctx) {
    rhg_attach_workspace(f.ctx, nullptr);
    rhg_use_stream(f.ctx, nullptr); // idle contexts hold no stream
  }
  return rc;
}

// One worker per slot: it takes the earliest arrived request whose framework
// is idle and fits the budget, runs it, and repeats until the loop stops
// with no request left.
void worker(rhg_server *s, int slot) {
  hipSetDevice(s->device);
  for (;;) {
    Job *jp = nullptr;
    FrameworkSpec *fp = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        if (s->stopping && s->waiting.empty()) return;
        const double now = secs(s->t0, Clock::now());
        double next_arrival = -1.0;
        bool blocked = false, progressed = false;
        for (auto it = s->waiting.begin(); it != s->waiting.end(); ++it) {
          Job &j = s->jobs[*it];
          if (j.arrival > now) {
            next_arrival = j.arrival;
            break;
          }
          FrameworkSpec &f = s->fw[j.req.framework];
          if (f.busy) continue;
          bool wait = false;
          const Clock::time_point ts = Clock::now();
          int rc = reserve(s, f, j.req.probe, slot, &wait);
          s->setup_s += secs(ts, Clock::now());
          if (wait && s->n_active > 0) {
            blocked = true; // memory frees when a running request ends
            break;
          }
          j.started = true;
          j.res.queued_s = j.arrival;
          j.res.started_s = now;
          s->waiting.erase(it);
          f.busy = true;
          progressed = true;
          if (rc != RHG_OK) {
            finish(s, j, f, rc);
            break;
          }
          s->ws_busy[slot] = 1;
          j.slot = slot;
          ++s->n_active;
          jp = &j;
          fp = &f;
          break;
        }
        if (jp) break;
        if (progressed) continue;
        // wake at the next arrival, or when a request is submitted or ends
        // (which frees a framework or memory) or the loop stops
        const Clock::time_point tw = Clock::now();
        if (!blocked && next_arrival >= 0.0)
          s->cv.wait_for(lk, std::chrono::duration<double>(std::max(next_arrival - now, 0.0) + 1e-5));
        else
          s->cv.wait(lk);
        s->busy_wait_s += secs(tw, Clock::now());
      }
    }
    double t_setup = 0, t_sub = 0, t_col = 0, work[12] = {0};
    bool created = false;
    int grids = 0;
    const int rc = execute(s, *jp, *fp, slot, &t_setup, &t_sub, &t_col, &created, &grids, work);
    std::lock_guard<std::mutex> lk(s->mu);
    s->work_ins += work[0];
    for (int k = 0; k < 3; ++k) s->work_pairs[k] += work[9 + k];
    s->setup_s += t_setup;
    s->submit_s += t_sub;
    s->collect_s += t_col;
    s->grids_built += grids;
    if (created && fp->ctx) s->n_created += 1;
    s->n_retries += jp->res.retries;
    --s->n_active;
    finish(s, *jp, *fp, rc);
  }
}

} // namespace

extern "C" {

int rhg_server_create(int device, size_t hbm_budget, int32_t slots, int32_t precision, int64_t slice,
                      rhg_server **out) {
  if (!out || slots < 1 || slots > 64 || slice < 1 || slice > ((int64_t)1 << 24)) return RHG_ERR_INVALID;
  if (precision != RHG_PREC_FP64 && precision != RHG_PREC_MIXED) return RHG_ERR_INVALID;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return RHG_ERR_HIP;
  rhg_server *s = new rhg_server();
  s->device = device;
  s->slots = slots;
  s->precision = precision;
  s->slice = slice;
  if (hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return RHG_ERR_HIP;
    }
    hbm_budget = (size_t)(0.8 * (double)fr);
  }
  s->budget = hbm_budget;
  s->pool = rhg_pool_create(device, (size_t)256 << 20);
  s->ws.assign((size_t)slots, nullptr);
  s->ws_bytes.assign((size_t)slots, 0);
  s->ws_busy.assign((size_t)slots, 0);
  s->streams.assign((size_t)slots, nullptr);
  for (int w = 0; w < slots; ++w) {
    rhg_workspace_create(device, &s->ws[w]);
    if (hipStreamCreateWithFlags(&s->streams[w], hipStreamNonBlocking) != hipSuccess) {
      rhg_server_destroy(s);
      return RHG_ERR_HIP;
    }
  }
  *out = s;
  return RHG_OK;
}

void rhg_server_destroy(rhg_server *s) {
  if (!s) return;
  rhg_server_stop(s);
  hipSetDevice(s->device);
  for (auto &kv : s->fw)
    if (kv.second.ctx) {
      rhg_destroy(kv.second.ctx);
      free_grids(s, kv.second);
    }
  for (rhg_workspace *w : s->ws) rhg_workspace_destroy(w);
  for (hipStream_t st : s->streams)
    if (st) hipStreamDestroy(st);
  rhg_pool_destroy(s->pool); // after every context that carved from it
  delete s;
}

const char *rhg_server_last_error(const rhg_server *s) { return s ? s->err.c_str() : "null server"; }

int rhg_server_add_framework(rhg_server *s, int64_t id, const rhg_framework *f, const rhg_cell *cell,
                             const rhg_forcefield *ff) {
  if (!s || !f || !cell || !ff || f->n_atoms < 0) return RHG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->fw.count(id)) return sfail(s, RHG_ERR_INVALID, "framework id already registered");
  FrameworkSpec &sp = s->fw[id];
  const int n = f->n_atoms;
  sp.x.assign(f->x, f->x + n);
  sp.y.assign(f->y, f->y + n);
  sp.z.assign(f->z, f->z + n);
  sp.eps.assign(f->eps_k, f->eps_k + n);
  sp.sig.assign(f->sigma, f->sigma + n);
  sp.q.assign((size_t)n, 0.0);
  if (f->charge) sp.q.assign(f->charge, f->charge + n);
  sp.mass.assign((size_t)n, 12.0);
  if (f->mass) sp.mass.assign(f->mass, f->mass + n);
  sp.cell = *cell;
  sp.ff = *ff;
  return RHG_OK;
}

int rhg_server_use_grids(rhg_server *s, int64_t id, double spacing) {
  if (!s || spacing < 0.0) return RHG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->fw.find(id);
  if (it == s->fw.end()) return sfail(s, RHG_ERR_INVALID, "unknown framework id");
  FrameworkSpec &f = it->second;
  if (f.busy) return sfail(s, RHG_ERR_INVALID, "framework in use");
  if (f.ctx) evict(s, f); // its context may hold grids of another spacing
  f.grid_spacing = spacing;
  return RHG_OK;
}

int rhg_server_set_tuning(rhg_server *s, int32_t n, const int32_t *max_atoms, const double *cell_size,
                          const int64_t *chunk, const int32_t *sites_per_lane, const int32_t *waves_per_block) {
  if (!s || n < 0 || (n > 0 && (!max_atoms || !cell_size || !chunk))) return RHG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<Tuning> rows;
  for (int i = 0; i < n; ++i) {
    const int32_t lanes = sites_per_lane ? sites_per_lane[i] : 4, waves = waves_per_block ? waves_per_block[i] : 2;
    if (!(cell_size[i] > 0.0) || chunk[i] < 1 || chunk[i] > ((int64_t)1 << 24) || lanes < 1 || lanes > 4 ||
        (waves != 1 && waves != 2 && waves != 4))
      return sfail(s, RHG_ERR_INVALID, "bad tuning row");
    rows.push_back({max_atoms[i], cell_size[i], chunk[i], lanes, waves});
  }
  std::sort(rows.begin(), rows.end(), [](const Tuning &a, const Tuning &b) { return a.max_atoms < b.max_atoms; });
  s->tuning = rows;
  return RHG_OK;
}

int rhg_server_submit(rhg_server *s, const rhg_request *r, int64_t *id) {
  if (!s || !r || !id) return RHG_ERR_INVALID;
  const rhg_probe &p = r->probe;
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->fw.count(r->framework)) return sfail(s, RHG_ERR_INVALID, "unknown framework id");
  if (p.n_sites < 1 || p.n_sites > RHG_MAX_SITES || p.start < 0 || p.start >= p.n_sites)
    return sfail(s, RHG_ERR_INVALID, "bad probe");
  if (r->n_insertions < 1 || r->steps_per_cycle < 1 || !(r->arrival_s >= 0.0))
    return sfail(s, RHG_ERR_INVALID, "bad request");
  Job j;
  memset(&j.res, 0, sizeof(j.res));
  j.req = *r;
  memset(&j.mol, 0, sizeof(j.mol));
  // a request submitted while the loop runs arrives now at the earliest
  j.arrival = s->running ? std::max(r->arrival_s, secs(s->t0, Clock::now())) : r->arrival_s;
  *id = (int64_t)s->jobs.size();
  s->jobs.push_back(j);
  auto pos = std::upper_bound(s->waiting.begin(), s->waiting.end(), j.arrival,
                              [&](double a, int64_t k) { return a < s->jobs[k].arrival; });
  s->waiting.insert(pos, *id);
  s->cv.notify_all();
  return RHG_OK;
}

int rhg_server_inject_fault(rhg_server *s, int64_t id, int64_t after_chunks) {
  if (!s) return RHG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (id < 0 || id >= (int64_t)s->jobs.size() || s->jobs[id].started) return RHG_ERR_INVALID;
  s->jobs[id].fault_after = after_chunks;
  return RHG_OK;
}

int rhg_server_start(rhg_server *s) {
  if (!s) return RHG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->running) return RHG_OK;
  s->running = true;
  s->stopping = false;
  s->t0 = Clock::now();
  for (int w = 0; w < s->slots; ++w) s->workers.emplace_back(worker, s, w);
  return RHG_OK;
}

int rhg_server_wait(rhg_server *s, int64_t id, double timeout_s) {
  if (!s) return RHG_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  if (id < 0 || id >= (int64_t)s->jobs.size()) return RHG_ERR_INVALID;
  auto done = [&]() { return s->jobs[id].finished; };
  if (timeout_s < 0.0)
    s->done_cv.wait(lk, done);
  else
    s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  return done() ? 1 : 0;
}

int rhg_server_stop(rhg_server *s) {
  if (!s) return RHG_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->running) return RHG_OK;
    s->stopping = true;
    s->cv.notify_all();
  }
  for (auto &t : s->workers) t.join();
  std::lock_guard<std::mutex> lk(s->mu);
  s->workers.clear();
  s->running = false;
  s->stopping = false;
  return RHG_OK;
}

int rhg_server_run(rhg_server *s) {
  if (int rc = rhg_server_start(s)) return rc;
  return rhg_server_stop(s);
}

int rhg_server_result(rhg_server *s, int64_t id, rhg_result *out) {
  if (!s || !out) return RHG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (id < 0 || id >= (int64_t)s->jobs.size()) return RHG_ERR_INVALID;
  *out = s->jobs[id].res;
  return s->jobs[id].finished ? RHG_OK : RHG_ERR_INVALID;
}

int rhg_server_stats(rhg_server *s, double out[20]) {
  if (!s || !out) return RHG_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  size_t gb = 0;
  for (auto &kv : s->fw) gb += kv.second.grid_bytes;
  out[0] = s->n_created;
  out[1] = s->n_evicted;
  out[2] = (double)s->peak_resident;
  out[3] = (double)s->resident;
  out[4] = s->n_retries;
  out[5] = (double)s->budget;
  out[6] = s->busy_wait_s;
  out[7] = (double)s->jobs.size();
  out[8] = s->setup_s;
  out[9] = s->submit_s;
  out[10] = s->collect_s;
  out[11] = s->grids_built;
  out[12] = s->grids_evicted;
  out[13] = (double)gb;
  out[14] = s->n_finished;
  out[15] = s->work_ins;
  out[16] = s->work_pairs[0];
  out[17] = s->work_pairs[1];
  out[18] = s->work_pairs[2];
  out[19] = 0.0;
  return RHG_OK;
}

size_t rhg_estimate_footprint(const rhg_framework *f, const rhg_cell *cell, const rhg_forcefield *ff,
                              int32_t n_types, int64_t chunk, int32_t n_sites, int32_t n_charged,
                              int32_t k_fb, int32_t k_or, double cell_size, int32_t precision) {
  if (!f || !cell || !ff || chunk < 1 || cell_size <= 0.0 || n_sites < 1) return 0;
  return estimate_bytes(f, cell, ff, n_types, chunk, n_sites, n_charged, k_fb, k_or, cell_size,
                        precision);
}

} // extern "C"
