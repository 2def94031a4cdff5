// C-ABI and host orchestration of the MI355X flow-solver hot path.
//
// The entry points are declared in include/flite3d_amd.h; each names the
// reference routine it replaces.  Kernels live in kernels.h.  One HIP stream
// per context; all calls return after the stream has drained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <queue>
#include <string>
#include <vector>
#include <cstdlib>
#include <cstdint>
#include <type_traits>

#include "kernels.h"

using namespace fk;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPTRY(x)                                                                  \
  do {                                                                             \
    hipError_t _e = (x);                                                           \
    if (_e != hipSuccess)                                                          \
      return fail(FLITE_E_HIP, std::string(#x) + ": " + hipGetErrorString(_e));    \
  } while (0)

#define TRY(x)              \
  do {                      \
    int _rc = (x);          \
    if (_rc != 0) return _rc; \
  } while (0)

inline int nblk(long long n, int t = NTHR) { return int((n + t - 1) / t); }

// device allocation list owned by a level
struct Pool {
  std::vector<void *> ptrs;
  ~Pool() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int alloc(T **out, size_t count) {
    void *p = nullptr;
    size_t bytes = sizeof(T) * (count ? count : 1);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(FLITE_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemset(p, 0, bytes);
    if (e != hipSuccess) return fail(FLITE_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    // the null-stream memset is asynchronous to the host for device memory
    // and unordered with the context's non-blocking stream: finish it before
    // a kernel there writes the buffer (lazily allocated scratch)
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(FLITE_E_HIP, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
    ptrs.push_back(p);
    *out = static_cast<T *>(p);
    return 0;
  }
  template <typename T>
  int upload(T **out, const std::vector<T> &v) {
    TRY(alloc(out, v.size()));
    if (!v.empty()) {
      hipError_t e = hipMemcpy(*out, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
      if (e != hipSuccess) return fail(FLITE_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    return 0;
  }
};

// Reverse Cuthill-McKee ordering of the side graph; returns host -> device
std::vector<int> rcm_order(int n, const std::vector<int> &e1, const std::vector<int> &e2) {
  std::vector<int> deg(n, 0);
  for (size_t e = 0; e < e1.size(); ++e) { deg[e1[e]]++; deg[e2[e]]++; }
  std::vector<int> ptr(n + 1, 0);
  for (int i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + deg[i];
  std::vector<int> adj(ptr[n]);
  std::vector<int> pos(ptr.begin(), ptr.end() - 1);
  for (size_t e = 0; e < e1.size(); ++e) { adj[pos[e1[e]]++] = e2[e]; adj[pos[e2[e]]++] = e1[e]; }
  for (int i = 0; i < n; ++i)
    std::sort(adj.begin() + ptr[i], adj.begin() + ptr[i + 1],
              [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> seen(n, 0);
  auto bfs = [&](int s, std::vector<int> *out, int *last) {
    std::vector<int> q{s};
    std::vector<char> vis(n, 0);
    vis[s] = 1;
    for (size_t h = 0; h < q.size(); ++h) {
      int v = q[h];
      for (int k = ptr[v]; k < ptr[v + 1]; ++k)
        if (!vis[adj[k]]) { vis[adj[k]] = 1; q.push_back(adj[k]); }
    }
    if (last) *last = q.back();
    if (out) *out = q;
  };
  for (int s0 = 0; s0 < n; ++s0) {
    if (seen[s0]) continue;
    // pseudo-peripheral start: two BFS sweeps from the component's first node
    int far = s0;
    bfs(s0, nullptr, &far);
    int start = far;
    std::vector<int> q{start};
    seen[start] = 1;
    for (size_t h = 0; h < q.size(); ++h) {
      int v = q[h];
      for (int k = ptr[v]; k < ptr[v + 1]; ++k)
        if (!seen[adj[k]]) { seen[adj[k]] = 1; q.push_back(adj[k]); }
    }
    order.insert(order.end(), q.begin(), q.end());
  }
  std::reverse(order.begin(), order.end());
  std::vector<int> h2d(n);
  for (int d = 0; d < n; ++d) h2d[order[d]] = d;
  return h2d;
}

}  // namespace

// host copy of one uploaded level view, 0-based, validated
struct Staged {
  int base = 0;  // index_base of the uploaded view
  int n = 0, ne = 0, nbs = 0, nb = 0, nfine = 0;
  std::vector<int> e1, e2, b1, b2, bi, nia, f2c;
  std::vector<double> w, wb, vol, xyz, wd, rw;  // w/wb/xyz column-major as in the view
  // moving mesh (empty on a fixed mesh): GCL weights per side and per
  // boundary side (Eb,2), previous volumes and movement (N,3) on level 1
  std::vector<double> gcl, bgcl, volp, volp2, cmv;
  int reg[21];
};

struct Level {
  bool fp32 = false;
  int n = 0, ne = 0, nbs = 0, nb = 0, gnum = 0, nfine = 0;
  int lanes = 1;  // lanes per node in the gather kernels (1, 4 or 16)
  int nbatch = 1;         // conditions sharing this mesh (state arrays are nbatch-fold)
  long long gstride = 0;  // gradient-record elements per condition
  bool has_coords = false;
  bool uploaded = false;
  std::vector<int> h2d, d2h;
  std::vector<int> wall_dev;  // wall nodes (indicator 1 range) in register order
  std::vector<double> warea_host;
  Pool pool;
  // device arrays (element type per precision, stored untyped)
  void *w, *wb, *vol, *xyz, *wd, *nrm, *warea;
  void *u, *uprev, *rhs, *diss, *src, *p, *dt, *dtt, *mu, *vort, *divg, *tdiff;
  void *rec0, *grec, *tau, *dnu, *fbuf, *bbuf, *ws;
  void *jst = nullptr;  // JST node records (fluxType 1-3 only)
  void *uold = nullptr, *uold2 = nullptr;  // unsteady only
  void *gcl = nullptr, *bgcl = nullptr, *volp = nullptr, *volp2 = nullptr, *cmv = nullptr;  // moving mesh only
  int *ea, *eb, *ba, *bb, *bind, *ell_ptr, *ell_k, *bs_ptr, *bs_list, *ncls, *host_index, *wall_nodes;
  int *ell_e, *ell_nbr;
  void *ellw;
  long long nslot = 0;
  // side blocks of the LDS-staged flux kernel, one table set per shape
  // (FB_SHAPE_THR / FB_SHAPE_NODES); fb_shape is the one in use
  struct SideBlocks {
    int nfb = 0;
    long long nent = 0;  // per-node partial-sum entries
    int *head = nullptr, *node = nullptr, *iptr = nullptr, *pn_ptr = nullptr, *pn_ent = nullptr;
    unsigned *loc = nullptr;
    unsigned short *inc = nullptr;
  } sb[FB_NSHAPES];
  int fb_shape = 0;
  void *pbuf = nullptr;  // sized for the largest nent of the shapes
  bool flux_partials = false;  // the last interior fluxes went to pbuf (k_flux_blk), not fbuf
  // node blocks of the LDS-staged gathers (ngb = 0: the per-node kernels)
  int ngb = 0;
  int *gb_first = nullptr, *gb_nptr = nullptr, *gb_node = nullptr;
  unsigned short *ell_loc = nullptr;
  // transfer from the next finer level (coarse levels)
  int *f2c = nullptr, *cptr = nullptr, *clist = nullptr;
  double *rw = nullptr;
  // monitor records: MON_BLOCKS block records and the total (MON_REC doubles each)
  double *m_part = nullptr;
  double *abuf = nullptr;  // adjoint side shares, 10 per side (allocated on first use)
  double *vmesh = nullptr;  // viscous mesh-adjoint shares, 3 per side and boundary side (first use)
  // second-order linearisation (allocated on first use): per-side adjoint
  // shares, per-node scratch, extremal neighbours of the limiter bounds
  double *abuf2 = nullptr, *lin = nullptr;
  int *amin = nullptr;
  double *vlin = nullptr;  // laminar viscous linearisation scratch (allocated on first use)
  std::vector<int> side_host;  // device side -> uploaded side index

  template <typename R>
  LevelDev<R> view() const {
    LevelDev<R> v;
    v.n = n; v.ne = ne; v.nbs = nbs; v.gnum = gnum;
    v.ea = ea; v.eb = eb; v.w = (const R *)w;
    v.ba = ba; v.bb = bb; v.bind = bind; v.wb = (const R *)wb;
    v.ell_ptr = ell_ptr; v.ell_k = ell_k; v.ell_e = ell_e; v.ell_nbr = ell_nbr; v.ellw = (const R *)ellw; v.nslot = nslot;
    v.bs_ptr = bs_ptr; v.bs_list = bs_list;
    const SideBlocks &B = sb[fb_shape];
    v.nfb = B.nfb; v.fb_head = B.head; v.fb_node = B.node; v.fb_loc = B.loc;
    v.nent = B.nent; v.fb_iptr = B.iptr; v.fb_inc = B.inc; v.pn_ptr = B.pn_ptr; v.pn_ent = B.pn_ent;
    v.pbuf = (R *)pbuf;
    v.ngb = ngb; v.gb_first = gb_first; v.gb_nptr = gb_nptr; v.gb_node = gb_node; v.ell_loc = ell_loc;
    v.vol = (const R *)vol; v.xyz = has_coords ? (const R *)xyz : nullptr;
    v.wd = (const R *)wd; v.nrm = (const R *)nrm; v.ncls = ncls;
    v.u = (R *)u; v.uprev = (R *)uprev; v.rhs = (R *)rhs; v.diss = (R *)diss; v.src = (R *)src;
    v.p = (R *)p; v.dt = (R *)dt; v.dtt = (R *)dtt; v.mu = (R *)mu; v.vort = (R *)vort;
    v.divg = (R *)divg; v.tdiff = (R *)tdiff; v.rec0 = (R *)rec0; v.grec = (R *)grec;
    v.tau = (R *)tau; v.dnu = (R *)dnu; v.fbuf = (R *)fbuf; v.bbuf = (R *)bbuf; v.ws = (R *)ws;
    v.jst = (R *)jst;
    v.uold = (R *)uold;
    v.uold2 = (R *)uold2;
    v.gcl = (const R *)gcl; v.bgcl = (const R *)bgcl;
    v.volp = (const R *)volp; v.volp2 = (const R *)volp2; v.cmv = (const R *)cmv;
    v.nb = nbatch;
    v.gstride = gstride;
    return v;
  }
};

struct flite_ctx {
  flite_params hp;
  DevParams P;
  int nl = 0, nbatch = 1;
  int lanes_override = 0;  // FLITE_GATHER_LANES: 1, 4 or 16 on every level (0 = by size)
  int shape_override = -1;  // FLITE_SIDE_BLOCK: side-block shape by sides per block (-1: shape 0)
  bool flux_per_side = false;  // FLITE_FLUX_KERNEL=side: one lane per side with direct record loads
  bool gather_per_node = false;  // FLITE_GATHER_KERNEL=node: no LDS-staged node blocks
  bool grad_direct = false;      // FLITE_GRAD_STORE=direct: per-thread record stores in the gradient
  // One cycle's launches form the same sequence with the same arguments
  // every cycle, so flite_cycle captures them once into a HIP graph and
  // replays it (FLITE_CYCLE_GRAPH=0: launch them one by one).  Anything that
  // changes the levels or the launch sequence drops the graph.
  bool graph_off = false;
  hipGraphExec_t cycle_exec = nullptr;
  std::vector<int> cycle_evals;  // evaluations per level counted while capturing
  std::vector<flite_params> hp_b;  // per condition (batched contexts)
  double *cond_dev = nullptr;      // per-condition Mach and inflow state on the device
  int sel = 0;                     // condition that field/state/monitor calls act on
  std::vector<std::unique_ptr<Level>> lev;
  // coarse levels of the other precision (flite_set_coarse_precision),
  // built on first use and swapped with lev[1..] on every switch
  std::vector<std::unique_ptr<Level>> lev_alt;
  std::vector<std::vector<int>> orders;  // device numbering per level
  hipStream_t stream = nullptr;
  // kernel timing (flite_timing_enable)
  std::string timed;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used;
  double timed_ms = 0;
  int timed_count = 0;
  // residual evaluations per level in the last counted cycle
  std::vector<int> evals;
  // uploaded level views, built on the device at first use
  std::vector<Staged> staged;
  std::vector<char> staged_ok;
  // solution arrays carried over a mesh replacement (flite_level_geometry):
  // per level u, uold, uold2 and laminar viscosity in host numbering, each
  // nbatch x ncomp blocks of n
  std::vector<std::vector<std::vector<double>>> carry;
  bool built = false;
  bool preamble_due = false;  // a replaced mesh still owes the step's preamble
  int device = 0;             // the HIP device current when the context was created
};

namespace {

template <typename F>
int with_level(const Level &L, F f) {
  if (L.fp32) return f(L.view<float>());
  return f(L.view<double>());
}

// f(std::integral_constant<int, SUB>) for the lanes-per-node variants built
template <typename F>
int with_lanes(int lanes, F f) {
  switch (lanes) {
    case 16: return f(std::integral_constant<int, 16>{});
    case 4: return f(std::integral_constant<int, 4>{});
    default: return f(std::integral_constant<int, 1>{});
  }
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FLITE_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// Device scratch of flite_monitor_noalloc: the MON_BLOCKS block records and
// the total, MON_REC doubles each (119 KB), held on every GPU the library
// runs on until the process ends.  In host code `scratch` names a host-side
// stand-in, not a device address: each GPU's copy is found once per device
// with hipGetSymbolAddress (scratch_setup, called by flite_ctx_create with
// that device current) and only that address is used.
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[SCRATCH_DEVICES] = {};
}  // namespace
static __device__ __align__(256) double scratch[MON_REC * (MON_BLOCKS + 1)];
namespace {

int scratch_setup(int dev) {
  if (dev < 0 || dev >= SCRATCH_DEVICES) return fail(FLITE_E_HIP, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev]) return 0;
  void *ptr = nullptr;
  HIPTRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail(FLITE_E_HIP, "no device address for the monitor scratch");
  g_scratch_dev[dev] = static_cast<double *>(ptr);
  return 0;
}

// one node walk feeding parts A, B, C.  Walks that read the neighbours'
// viscous records run on the LDS-staged node blocks where the level has
// them; walks that read only the 64-byte gather records stay one thread
// group of L.lanes lanes per node, which measured faster for them.
template <typename A, typename B, typename C, typename V>
void launch_gather(flite_ctx *c, const Level &L, const V &v, const DevParams &P) {
  using R = std::remove_pointer_t<decltype(v.u)>;
  if (L.ngb > 0 && L.lanes <= 4 && (A::AUX || B::AUX || C::AUX)) {
    hipLaunchKernelGGL((k_gather_blk<R, A, B, C>), dim3((unsigned)L.ngb * L.nbatch), dim3(NTHR), 0, c->stream, v, P);
    return;
  }
  with_lanes(L.lanes, [&](auto sc) {
    constexpr int SUB = decltype(sc)::value;
    hipLaunchKernelGGL((k_gather<R, SUB, A, B, C>), dim3(nblk((long long)L.n * SUB) * L.nbatch), dim3(NTHR), 0,
                       c->stream, v, P);
    return 0;
  });
}

int sync(flite_ctx *c) {
  HIPTRY(hipStreamSynchronize(c->stream));
  return 0;
}

hipEvent_t take_event(flite_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// scoped timing of one launch of kernel class `name` on level `gnum`
struct Timed {
  flite_ctx *c;
  bool on;
  hipEvent_t a = nullptr, b = nullptr;
  Timed(flite_ctx *c_, const char *name, int gnum) : c(c_) {
    on = !c->timed.empty() && c->timed == std::string(name) + "_l" + std::to_string(gnum);
    if (on) {
      a = take_event(c);
      b = take_event(c);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~Timed() {
    if (on) {
      (void)hipEventRecord(b, c->stream);
      c->ev_used.push_back({a, b});
    }
  }
};

int harvest_timing(flite_ctx *c) {
  HIPTRY(hipStreamSynchronize(c->stream));
  for (auto &pr : c->ev_used) {
    float ms = 0;
    HIPTRY(hipEventElapsedTime(&ms, pr.first, pr.second));
    c->timed_ms += ms;
    c->timed_count += 1;
    c->ev_pool.push_back(pr.first);
    c->ev_pool.push_back(pr.second);
  }
  c->ev_used.clear();
  return 0;
}

DevParams make_dev_params(const flite_params &p) {
  DevParams P;
  P.gamma = p.gamma;
  P.mach = p.mach_number;
  P.Re = p.reynolds_number;
  P.Pr = p.prandtl_number;
  P.Prt = p.turbulent_prandtl_number;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 6; ++j) P.inflow[k][j] = p.inflow_field[k][j];
  P.inflow_temperature = p.inflow_temperature;
  P.isentropic = p.walls_are_isentropic;
  P.sa = p.turbulence_model == 1;
  P.order = p.hllc_flux_order;
  P.autoclip = p.automatic_clipping;
  P.nsig = p.reynolds_number < 1 ? 4 : 2;
  P.bterm = p.boundary_term;
  P.clip = p.hllc_clipping;
  P.max_inc = p.max_increment_factor;
  P.min_rho = p.minimum_density;
  P.max_turb = p.max_turbulence_value;
  P.turb_k = p.turb_k;
  P.beta0 = std::sqrt((p.gamma - 1.0) / (2.0 * p.gamma));
  P.prolong_relax = p.prolongation_relaxation;
  P.turb_prolong_relax = p.turb_prolongation_relaxation;
  P.mg_bc_relax = p.multigrid_bc_relaxation;
  P.ftype = p.flux_type;
  P.matdiss = p.use_matrix_dissipation;
  P.k2 = p.second_order_dissipation_factor;
  P.k4 = p.fourth_order_dissipation_factor;
  P.k4c = p.coarse_grid_dissipation_factor;
  P.hcf = p.hartens_correction_factor;
  P.unsteady = p.flow_type != 0;
  P.add_time = p.add_time_step;
  P.bdf2 = p.time_scheme == 2;
  P.dtp = p.physical_time_step;
  return P;
}

// ---------------------------------------------------------------------------
// level upload
// ---------------------------------------------------------------------------
template <typename R>
int upload_real(Pool &pool, void **out, const std::vector<double> &v) {
  std::vector<R> t(v.begin(), v.end());
  R *d;
  TRY(pool.upload(&d, t));
  *out = d;
  return 0;
}

template <typename R>
int alloc_real(Pool &pool, void **out, size_t count) {
  R *d;
  TRY(pool.alloc(&d, count));
  *out = d;
  return 0;
}

int stage_level(flite_ctx *c, int level, const flite_level_view *v, Staged &S) {
  if (!v) return fail(FLITE_E_ARG, "null level view");
  const int b = v->index_base;
  if (b != 0 && b != 1) return fail(FLITE_E_ARG, "index_base must be 0 or 1");
  S.base = b;
  S.n = v->n_nodes;
  S.ne = v->n_sides;
  S.nbs = v->n_bsides;
  S.nb = v->n_bnodes;
  const int n = S.n, ne = S.ne, nbs = S.nbs;
  if (n <= 0 || ne < 0 || nbs < 0 || S.nb < 0 || S.nb > n) return fail(FLITE_E_ARG, "bad level sizes");
  if (!v->side_index || !v->side_weights || !v->bside_index || !v->bside_weights || !v->node_volume ||
      !v->wall_distance || !v->node_indicator_register || !v->node_indicator_array)
    return fail(FLITE_E_ARG, "missing level array");
  S.e1.resize(ne);
  S.e2.resize(ne);
  for (int e = 0; e < ne; ++e) {
    S.e1[e] = v->side_index[e] - b;
    S.e2[e] = v->side_index[(size_t)ne + e] - b;
    if (S.e1[e] < 0 || S.e1[e] >= n || S.e2[e] < 0 || S.e2[e] >= n) return fail(FLITE_E_ARG, "side index out of range");
  }
  S.w.assign(v->side_weights, v->side_weights + 3 * (size_t)ne);
  S.b1.resize(nbs);
  S.b2.resize(nbs);
  S.bi.resize(nbs);
  for (int s = 0; s < nbs; ++s) {
    S.b1[s] = v->bside_index[s] - b;
    S.b2[s] = v->bside_index[(size_t)nbs + s] - b;
    S.bi[s] = v->bside_index[2 * (size_t)nbs + s];
    if (S.b1[s] < 0 || S.b1[s] >= n || S.b2[s] < 0 || S.b2[s] >= n)
      return fail(FLITE_E_ARG, "boundary side index out of range");
    if (!(S.bi[s] <= 4 || S.bi[s] == 9)) return fail(FLITE_E_UNSUPPORTED, "engine boundary sides are not supported");
  }
  S.wb.assign(v->bside_weights, v->bside_weights + 3 * (size_t)nbs);
  S.vol.assign(v->node_volume, v->node_volume + n);
  for (int i = 0; i < n; ++i)
    if (!(S.vol[i] > 0)) return fail(FLITE_E_ARG, "non-positive node volume");
  S.wd.assign(v->wall_distance, v->wall_distance + n);
  if (level == 1 && v->coordinates) S.xyz.assign(v->coordinates, v->coordinates + 3 * (size_t)n);
  for (int k = 0; k < 21; ++k) S.reg[k] = v->node_indicator_register[k];
  S.nia.resize(S.nb);
  for (int i = 0; i < S.nb; ++i) {
    S.nia[i] = v->node_indicator_array[i] - b;
    if (S.nia[i] < 0 || S.nia[i] >= n) return fail(FLITE_E_ARG, "node indicator array out of range");
  }
  for (int k = 1; k <= 10; ++k) {
    int st = S.reg[2 * k - 2], en = S.reg[2 * k - 1];
    if (en < st) continue;
    if (st < 1 || en > S.nb) return fail(FLITE_E_ARG, "node indicator register out of range");
    if (k >= 5 && k <= 8) return fail(FLITE_E_UNSUPPORTED, "engine boundary nodes are not supported");
    if (k == 10) return fail(FLITE_E_UNSUPPORTED, "indicator 10 nodes are not supported");
  }
  if (level > 1) {
    if (!v->prolongation || !v->restriction) return fail(FLITE_E_ARG, "coarse level needs prolongation/restriction");
    S.nfine = v->n_fine;
    if (S.nfine <= 0) return fail(FLITE_E_ARG, "bad n_fine");
    S.f2c.resize(S.nfine);
    for (int i = 0; i < S.nfine; ++i) {
      S.f2c[i] = v->prolongation[i] - b;
      if (S.f2c[i] < 0 || S.f2c[i] >= n) return fail(FLITE_E_ARG, "prolongation index out of range");
    }
    S.rw.assign(v->restriction, v->restriction + S.nfine);
  }
  // moving-mesh arrays: used on unsteady runs only (the physical time step
  // divides them); each may be absent
  if (c->hp.flow_type != 0) {
    if (v->side_gcl) S.gcl.assign(v->side_gcl, v->side_gcl + ne);
    if (v->bside_gcl) S.bgcl.assign(v->bside_gcl, v->bside_gcl + 2 * (size_t)nbs);
    if (level == 1) {
      if (v->node_volume_p) S.volp.assign(v->node_volume_p, v->node_volume_p + n);
      if (v->node_volume_p2) S.volp2.assign(v->node_volume_p2, v->node_volume_p2 + n);
      if (v->coordinate_movement) S.cmv.assign(v->coordinate_movement, v->coordinate_movement + 3 * (size_t)n);
    }
  }
  return 0;
}

// Device numbering: RCM on the coarsest level; every finer level lists the
// members of each coarse node contiguously, in the coarse order, so that
// restriction and prolongation are contiguous-range operations and each
// agglomerate's nodes share cache lines.
// Z-order (Morton) of points: each axis is first replaced by its rank among
// the points, so that graded spacings (a boundary layer) still spread over
// all the curve's cells; ties keep the input order.  Returns host -> device.
std::vector<int> morton_order(const std::vector<double> &x, const std::vector<double> &y,
                              const std::vector<double> &z) {
  int n = (int)x.size();
  auto ranks = [n](const std::vector<double> &v) {
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] < v[b]; });
    std::vector<uint32_t> r(n);
    for (int k = 0; k < n; ++k) r[idx[k]] = (uint32_t)((uint64_t)k * ((1u << 21) - 1) / std::max(1, n - 1));
    return r;
  };
  auto spread = [](uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  std::vector<uint32_t> rx = ranks(x), ry = ranks(y), rz = ranks(z);
  std::vector<uint64_t> key(n);
  for (int i = 0; i < n; ++i) key[i] = spread(rx[i]) | spread(ry[i]) << 1 | spread(rz[i]) << 2;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
  std::vector<int> h2d(n);
  for (int d = 0; d < n; ++d) h2d[idx[d]] = d;
  return h2d;
}

// order of the coarsest level: Z-order of the centroids of its level-1
// members when level 1 has coordinates (FLITE_NODE_ORDER=rcm selects reverse
// Cuthill-McKee of the coarsest side graph instead)
std::vector<int> coarsest_order(const std::vector<Staged> &st) {
  int nl = (int)st.size();
  const Staged &F = st[0], &C = st[nl - 1];
  const char *e = std::getenv("FLITE_NODE_ORDER");
  bool rcm = e && std::string(e) == "rcm";
  if (rcm || F.xyz.empty()) return rcm_order(C.n, C.e1, C.e2);
  std::vector<int> m(F.n);
  std::iota(m.begin(), m.end(), 0);
  for (int l = 1; l < nl; ++l)
    for (int i = 0; i < F.n; ++i) m[i] = st[l].f2c[m[i]];
  std::vector<double> x(C.n, 0.0), y(C.n, 0.0), z(C.n, 0.0), cnt(C.n, 0.0);
  for (int i = 0; i < F.n; ++i) {
    x[m[i]] += F.xyz[i];
    y[m[i]] += F.xyz[(size_t)F.n + i];
    z[m[i]] += F.xyz[2 * (size_t)F.n + i];
    cnt[m[i]] += 1.0;
  }
  for (int k = 0; k < C.n; ++k) {
    double w = cnt[k] > 0 ? 1.0 / cnt[k] : 0.0;
    x[k] *= w;
    y[k] *= w;
    z[k] *= w;
  }
  return morton_order(x, y, z);
}

std::vector<std::vector<int>> nested_orders(const std::vector<Staged> &st) {
  int nl = (int)st.size();
  std::vector<std::vector<int>> h2d(nl);
  h2d[nl - 1] = coarsest_order(st);
  for (int l = nl - 2; l >= 0; --l) {
    const Staged &K = st[l + 1];
    int nc = K.n, nf = st[l].n;
    std::vector<int> cd2h(nc);
    for (int h = 0; h < nc; ++h) cd2h[h2d[l + 1][h]] = h;
    std::vector<int> cnt(nc + 1, 0);
    for (int i = 0; i < nf; ++i) cnt[K.f2c[i] + 1]++;
    for (int k = 0; k < nc; ++k) cnt[k + 1] += cnt[k];
    std::vector<int> members(nf);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < nf; ++i) members[pos[K.f2c[i]]++] = i;
    h2d[l].assign(nf, 0);
    int d = 0;
    for (int dc = 0; dc < nc; ++dc) {
      int hc = cd2h[dc];
      for (int q = cnt[hc]; q < cnt[hc + 1]; ++q) h2d[l][members[q]] = d++;
    }
  }
  return h2d;
}

// Side blocks of one shape for the LDS-staged flux: consecutive sides, at
// most FB_SHAPE_THR[shape] of them touching at most FB_SHAPE_NODES[shape]
// distinct nodes; each block's node list is ascending (neighbouring records
// share cache lines)
int build_side_blocks(Level &L, int shape, const std::vector<int> &ea, const std::vector<int> &eb) {
  const int n = L.n, ne = L.ne, FB_SIDES = FB_SHAPE_THR[shape], FB_NODES = FB_SHAPE_NODES[shape];
  Level::SideBlocks &B = L.sb[shape];
  std::vector<int> fb_head{0, 0, 0, 0}, fb_node, fb_nodep, fb_iptr{0};
  std::vector<unsigned> fb_loc(ne);
  std::vector<unsigned short> fb_inc;
  std::vector<int> stamp(n, -1), local(n, 0), cur;
  std::vector<std::vector<unsigned short>> inc;
  int start = 0, blk = 0;
  auto close = [&](int end) {
    std::sort(cur.begin(), cur.end());
    for (size_t q = 0; q < cur.size(); ++q) local[cur[q]] = (int)q;
    for (int k = start; k < end; ++k) fb_loc[k] = (unsigned)local[ea[k]] | ((unsigned)local[eb[k]] << 16);
    // each node's sides within the block, in side order
    inc.assign(cur.size(), {});
    for (int k = start; k < end; ++k) {
      inc[local[ea[k]]].push_back((unsigned short)((k - start) << 1));
      inc[local[eb[k]]].push_back((unsigned short)(((k - start) << 1) | 1));
    }
    for (auto &l : inc) {
      fb_inc.insert(fb_inc.end(), l.begin(), l.end());
      fb_iptr.push_back((int)fb_inc.size());
    }
    fb_node.insert(fb_node.end(), cur.begin(), cur.end());
    fb_nodep.insert(fb_nodep.end(), cur.begin(), cur.end());
    fb_nodep.resize((size_t)(blk + 1) * FB_NODES, -1);
    for (int v : {end, (int)fb_node.size(), fb_iptr.back(), 0}) fb_head.push_back(v);
    cur.clear();
    start = end;
    ++blk;
  };
  for (int k = 0; k < ne; ++k) {
    int add = (stamp[ea[k]] != blk) + (stamp[eb[k]] != blk && eb[k] != ea[k]);
    if (k - start == FB_SIDES || (int)cur.size() + add > FB_NODES) close(k);
    for (int v : {ea[k], eb[k]})
      if (stamp[v] != blk) { stamp[v] = blk; cur.push_back(v); }
  }
  if (start < ne) close(ne);
  B.nfb = blk;
  B.nent = (long long)fb_node.size();
  // node -> its entries, in block order
  std::vector<int> pn_ptr(n + 1, 0), pn_ent(fb_node.size());
  for (int g : fb_node) pn_ptr[g + 1]++;
  for (int i = 0; i < n; ++i) pn_ptr[i + 1] += pn_ptr[i];
  {
    std::vector<int> pos(pn_ptr.begin(), pn_ptr.end() - 1);
    for (size_t q = 0; q < fb_node.size(); ++q) pn_ent[pos[fb_node[q]]++] = (int)q;
  }
  TRY(L.pool.upload(&B.head, fb_head));
  TRY(L.pool.upload(&B.node, fb_nodep));
  TRY(L.pool.upload(&B.loc, fb_loc));
  TRY(L.pool.upload(&B.iptr, fb_iptr));
  TRY(L.pool.upload(&B.inc, fb_inc));
  TRY(L.pool.upload(&B.pn_ptr, pn_ptr));
  TRY(L.pool.upload(&B.pn_ent, pn_ent));
  return 0;
}

int build_level(flite_ctx *c, int level, const Staged &S, const std::vector<int> &h2d, bool coarse_fp32,
                std::unique_ptr<Level> &out) {
  auto Lp = std::make_unique<Level>();
  Level &L = *Lp;
  L.gnum = level;
  L.fp32 = coarse_fp32 && level > 1;
  L.n = S.n;
  // enough lanes to give the chip >= ~1M threads on mid-size levels, and
  // short per-lane gather chains on the small coarse ones
  L.nbatch = c->nbatch;
  L.lanes = c->lanes_override ? c->lanes_override : (S.n >= (1 << 20) ? 1 : S.n >= (1 << 16) ? 4 : 16);
  L.fb_shape = c->shape_override >= 0 ? c->shape_override : 0;
  L.ne = S.ne;
  L.nbs = S.nbs;
  L.nb = S.nb;
  L.nfine = S.nfine;
  const int n = L.n, ne = L.ne, nbs = L.nbs;
  L.h2d = h2d;
  L.d2h.assign(n, 0);
  for (int h = 0; h < n; ++h) L.d2h[L.h2d[h]] = h;
  // node classes from nodeIndicatorArray ranges
  std::vector<int> ncls(n, 0);
  for (int k = 1; k <= 10; ++k) {
    int st = S.reg[2 * k - 2], en = S.reg[2 * k - 1];
    for (int ib = st; ib <= en; ++ib) {
      int h = S.nia[ib - 1];
      ncls[L.h2d[h]] = k;
      if (k == 1) L.wall_dev.push_back(L.h2d[h]);
    }
  }
  for (int h = 0; h < L.nb; ++h) ncls[L.h2d[h]] |= NC_BFIRST;
  // normals (makeBoundaryNormals, DEF:15394-15460) and wall area
  // (makeSurfaceArea, DEF:15462-15540) in host numbering
  std::vector<double> nrm_h(3 * (size_t)n, 0.0), warea_h(n, 0.0);
  for (int s = 0; s < nbs; ++s)
    for (int k = 0; k < 3; ++k) {
      double w = S.wb[(size_t)k * nbs + s];
      nrm_h[3 * (size_t)S.b1[s] + k] -= w;
      nrm_h[3 * (size_t)S.b2[s] + k] -= w;
    }
  for (int i = 0; i < L.nb; ++i) {
    double x = nrm_h[3 * i], y = nrm_h[3 * i + 1], z = nrm_h[3 * i + 2];
    double nn = std::sqrt(x * x + y * y + z * z);
    for (int k = 0; k < 3; ++k) nrm_h[3 * i + k] = nn > 1e-10 ? nrm_h[3 * i + k] / nn : 0.0;
  }
  if (level == 1)
    for (int s = 0; s < nbs; ++s) {
      double wx = S.wb[s], wy = S.wb[nbs + s], wz = S.wb[2 * (size_t)nbs + s];
      double a = std::sqrt(wx * wx + wy * wy + wz * wz);
      warea_h[S.b1[s]] += 4.0 * a;
      warea_h[S.b2[s]] += 4.0 * a;
    }
  // device edges sorted by their lower device end
  std::vector<int> eorder(ne);
  std::iota(eorder.begin(), eorder.end(), 0);
  std::vector<int> da(ne), db(ne);
  for (int e = 0; e < ne; ++e) { da[e] = L.h2d[S.e1[e]]; db[e] = L.h2d[S.e2[e]]; }
  std::sort(eorder.begin(), eorder.end(), [&](int x, int y) {
    int lx = std::min(da[x], db[x]), ly = std::min(da[y], db[y]);
    if (lx != ly) return lx < ly;
    return std::max(da[x], db[x]) < std::max(da[y], db[y]);
  });
  L.side_host = eorder;
  std::vector<int> ea(ne), eb(ne);
  std::vector<double> w(3 * (size_t)ne);
  for (int k = 0; k < ne; ++k) {
    int e = eorder[k];
    ea[k] = da[e];
    eb[k] = db[e];
    for (int q = 0; q < 3; ++q) w[3 * (size_t)k + q] = S.w[(size_t)q * ne + e];
  }
  for (int k = 0; k < FB_NSHAPES; ++k) TRY(build_side_blocks(L, k, ea, eb));
  // sliced ELL incidence
  std::vector<int> deg(n, 0);
  for (int k = 0; k < ne; ++k) { deg[ea[k]]++; deg[eb[k]]++; }
  int nslice = (n + SLICE - 1) / SLICE;
  std::vector<int> ell_ptr(nslice + 1, 0), ell_k(nslice, 0);
  for (int s = 0; s < nslice; ++s) {
    int K = 0;
    for (int i = s * SLICE; i < std::min(n, (s + 1) * SLICE); ++i) K = std::max(K, deg[i]);
    ell_k[s] = K;
    ell_ptr[s + 1] = ell_ptr[s] + K * SLICE;
  }
  const size_t nslot = ell_ptr[nslice] ? ell_ptr[nslice] : 1;
  std::vector<int> ell_e(nslot, -1), ell_nbr(nslot, -1);
  std::vector<double> ellw(3 * nslot, 0.0);
  std::vector<int> fill(n, 0);
  for (int k = 0; k < ne; ++k) {
    for (int r = 0; r < 2; ++r) {
      int i = r ? eb[k] : ea[k], j = r ? ea[k] : eb[k];
      int s = i / SLICE, lane = i % SLICE;
      size_t slot = ell_ptr[s] + fill[i] * SLICE + lane;
      ell_e[slot] = (k << 1) | r;
      ell_nbr[slot] = j;
      for (int q = 0; q < 3; ++q) ellw[q * nslot + slot] = r ? -w[3 * (size_t)k + q] : w[3 * (size_t)k + q];
      fill[i]++;
    }
  }
  L.nslot = (long long)nslot;
  // node blocks for the LDS-staged gathers (used while the level runs one or
  // four lanes per node, when they give enough blocks to fill the chip):
  // consecutive nodes, at most GB_NODES of them, whose neighbours are at
  // most GB_CAP distinct nodes; each block's neighbour list is ascending
  if (n > 0 && !c->gather_per_node && *std::max_element(deg.begin(), deg.end()) <= GB_CAP) {
    std::vector<int> gb_first{0}, gb_nptr{0}, gb_node;
    std::vector<unsigned short> ell_loc(nslot, 0xffff);
    std::vector<int> stamp(n, -1), seen(n, -1), local(n, 0), cur;
    int start = 0, blk = 0;
    auto nbr_of = [&](int i, int k) {
      return ell_nbr[ell_ptr[i / SLICE] + (size_t)k * SLICE + i % SLICE];
    };
    auto close = [&](int end) {
      std::sort(cur.begin(), cur.end());
      for (size_t q = 0; q < cur.size(); ++q) local[cur[q]] = (int)q;
      for (int i = start; i < end; ++i)
        for (int k = 0; k < ell_k[i / SLICE]; ++k) {
          size_t slot = ell_ptr[i / SLICE] + (size_t)k * SLICE + i % SLICE;
          if (ell_nbr[slot] >= 0) ell_loc[slot] = (unsigned short)local[ell_nbr[slot]];
        }
      gb_node.insert(gb_node.end(), cur.begin(), cur.end());
      gb_nptr.push_back((int)gb_node.size());
      gb_first.push_back(end);
      cur.clear();
      start = end;
      ++blk;
    };
    for (int i = 0; i < n; ++i) {
      int add = 0;
      for (int k = 0; k < deg[i]; ++k) {
        int j = nbr_of(i, k);
        if (stamp[j] != blk && seen[j] != i) { seen[j] = i; ++add; }
      }
      if (i - start == GB_NODES || (int)cur.size() + add > GB_CAP) close(i);
      for (int k = 0; k < deg[i]; ++k) {
        int j = nbr_of(i, k);
        if (stamp[j] != blk) { stamp[j] = blk; cur.push_back(j); }
      }
    }
    if (start < n) close(n);
    L.ngb = blk;
    TRY(L.pool.upload(&L.gb_first, gb_first));
    TRY(L.pool.upload(&L.gb_nptr, gb_nptr));
    TRY(L.pool.upload(&L.gb_node, gb_node));
    TRY(L.pool.upload(&L.ell_loc, ell_loc));
  }
  // boundary sides in device numbering, per-node CSR
  std::vector<int> ba(nbs), bb(nbs);
  std::vector<double> wb(3 * (size_t)nbs);
  for (int s = 0; s < nbs; ++s) {
    ba[s] = L.h2d[S.b1[s]];
    bb[s] = L.h2d[S.b2[s]];
    for (int q = 0; q < 3; ++q) wb[3 * (size_t)s + q] = S.wb[(size_t)q * nbs + s];
  }
  std::vector<int> bs_ptr(n + 1, 0);
  for (int s = 0; s < nbs; ++s) { bs_ptr[ba[s] + 1]++; bs_ptr[bb[s] + 1]++; }
  for (int i = 0; i < n; ++i) bs_ptr[i + 1] += bs_ptr[i];
  std::vector<int> bs_list(bs_ptr[n] ? bs_ptr[n] : 1);
  {
    std::vector<int> pos(bs_ptr.begin(), bs_ptr.end() - 1);
    for (int s = 0; s < nbs; ++s) {
      bs_list[pos[ba[s]]++] = (s << 1) | 0;
      bs_list[pos[bb[s]]++] = (s << 1) | 1;
    }
  }
  // node arrays in device numbering
  std::vector<double> vol(n), wd(n), nrm(3 * (size_t)n), warea(n), xyz;
  for (int h = 0; h < n; ++h) {
    int d = L.h2d[h];
    vol[d] = S.vol[h];
    wd[d] = S.wd[h];
    warea[d] = warea_h[h];
    for (int k = 0; k < 3; ++k) nrm[3 * (size_t)d + k] = nrm_h[3 * (size_t)h + k];
  }
  L.warea_host = warea;
  L.has_coords = level == 1 && !S.xyz.empty();
  if (L.has_coords) {
    xyz.resize(3 * (size_t)n);
    for (int h = 0; h < n; ++h)
      for (int k = 0; k < 3; ++k) xyz[3 * (size_t)L.h2d[h] + k] = S.xyz[(size_t)k * n + h];
  }
  Pool &P = L.pool;
  TRY(P.upload(&L.ea, ea));
  TRY(P.upload(&L.eb, eb));
  TRY(P.upload(&L.ba, ba));
  TRY(P.upload(&L.bb, bb));
  TRY(P.upload(&L.bind, S.bi));
  TRY(P.upload(&L.ell_ptr, ell_ptr));
  TRY(P.upload(&L.ell_k, ell_k));
  TRY(P.upload(&L.ell_e, ell_e));
  TRY(P.upload(&L.ell_nbr, ell_nbr));
  TRY(P.upload(&L.bs_ptr, bs_ptr));
  TRY(P.upload(&L.bs_list, bs_list));
  TRY(P.upload(&L.ncls, ncls));
  TRY(P.upload(&L.host_index, L.d2h));
  TRY(P.upload(&L.wall_nodes, L.wall_dev));
  auto up_real = [&](void **out, const std::vector<double> &vv) {
    return L.fp32 ? upload_real<float>(P, out, vv) : upload_real<double>(P, out, vv);
  };
  auto al_real = [&](void **out, size_t cnt) {
    return L.fp32 ? alloc_real<float>(P, out, cnt) : alloc_real<double>(P, out, cnt);
  };
  const size_t NB = (size_t)L.nbatch;
  TRY(up_real(&L.w, w));
  TRY(up_real(&L.ellw, ellw));
  TRY(up_real(&L.wb, wb));
  TRY(up_real(&L.vol, vol));
  TRY(up_real(&L.wd, wd));
  TRY(up_real(&L.nrm, nrm));
  TRY(up_real(&L.warea, warea));
  if (L.has_coords) TRY(up_real(&L.xyz, xyz)); else L.xyz = nullptr;
  // moving mesh: the same permutations as w, wb and vol
  if (!S.gcl.empty()) {
    std::vector<double> g(ne);
    for (int k = 0; k < ne; ++k) g[k] = S.gcl[L.side_host[k]];
    TRY(up_real(&L.gcl, g));
  }
  if (!S.bgcl.empty()) {
    std::vector<double> g(2 * (size_t)nbs);
    for (int s = 0; s < nbs; ++s) {
      g[2 * (size_t)s] = S.bgcl[s];
      g[2 * (size_t)s + 1] = S.bgcl[(size_t)nbs + s];
    }
    TRY(up_real(&L.bgcl, g));
  }
  auto up_nodes = [&](void **out, const std::vector<double> &hv, int nc) -> int {
    if (hv.empty()) return 0;
    std::vector<double> d((size_t)nc * n);
    for (int h = 0; h < n; ++h)
      for (int k = 0; k < nc; ++k) d[(size_t)nc * L.h2d[h] + k] = hv[(size_t)k * n + h];
    return up_real(out, d);
  };
  TRY(up_nodes(&L.volp, S.volp, 1));
  TRY(up_nodes(&L.volp2, S.volp2, 1));
  TRY(up_nodes(&L.cmv, S.cmv, 3));
  TRY(al_real(&L.u, (size_t)(6 * (size_t)n) * NB));
  TRY(al_real(&L.uprev, (size_t)(6 * (size_t)n) * NB));
  TRY(al_real(&L.rhs, (size_t)(6 * (size_t)n) * NB));
  TRY(al_real(&L.diss, (size_t)(5 * (size_t)n) * NB));
  TRY(al_real(&L.src, (size_t)(6 * (size_t)n) * NB));
  TRY(al_real(&L.p, (size_t)(n) * NB));
  TRY(al_real(&L.dt, (size_t)(n) * NB));
  TRY(al_real(&L.dtt, (size_t)(n) * NB));
  TRY(al_real(&L.mu, (size_t)(n) * NB));
  TRY(al_real(&L.vort, (size_t)(n) * NB));
  TRY(al_real(&L.divg, (size_t)(n) * NB));
  TRY(al_real(&L.tdiff, (size_t)(n) * NB));
  TRY(al_real(&L.rec0, (size_t)(8 * (size_t)n) * NB));
  bool need_grad = L.has_coords && c->hp.hllc_flux_order > 1;
  L.gstride = need_grad ? GREC * (long long)n : 0;
  TRY(al_real(&L.grec, need_grad ? GREC * (size_t)n * NB : 1));
  TRY(al_real(&L.tau, (size_t)(VREC * (size_t)n) * NB));  // tau and dnu share the record
  L.dnu = (char *)L.tau + 9 * (L.fp32 ? 4 : 8);
  TRY(al_real(&L.fbuf, (size_t)(6 * (size_t)ne) * NB));
  {
    long long nent = 0;
    for (const auto &B : L.sb) nent = std::max(nent, B.nent);
    TRY(al_real(&L.pbuf, (size_t)(6 * (size_t)nent) * NB));
  }
  TRY(al_real(&L.bbuf, (size_t)(10 * (size_t)nbs) * NB));
  TRY(al_real(&L.ws, (size_t)(3 * (size_t)n) * NB));
  TRY(al_real(&L.jst, c->hp.flux_type != 4 ? (size_t)(8 * (size_t)n) * NB : 1));
  TRY(al_real(&L.uold, c->hp.flow_type != 0 ? (size_t)(6 * (size_t)n) * NB : 1));
  TRY(al_real(&L.uold2, c->hp.flow_type != 0 ? (size_t)(6 * (size_t)n) * NB : 1));
  TRY(P.alloc(&L.m_part, (size_t)MON_REC * (MON_BLOCKS + 1)));
  // transfer operators from the finer level
  if (level > 1) {
    Level &F = *c->lev[level - 2];
    std::vector<int> f2c(F.n);
    std::vector<double> rw(F.n);
    std::vector<int> cnt(n + 1, 0);
    for (int hf = 0; hf < F.n; ++hf) {
      int hc = S.f2c[hf];
      f2c[F.h2d[hf]] = L.h2d[hc];
      rw[F.h2d[hf]] = S.rw[hf];
      cnt[L.h2d[hc] + 1]++;
    }
    for (int k = 0; k < n; ++k) cnt[k + 1] += cnt[k];
    std::vector<int> clist(F.n);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int df = 0; df < F.n; ++df) clist[pos[f2c[df]]++] = df;  // fine members in device order
    for (int df = 0; df < F.n; ++df)  // nested order: members of each coarse node are contiguous
      if (clist[df] != df) return fail(FLITE_E_STATE, "fine level is not in nested order of its coarse level");
    TRY(P.upload(&L.f2c, f2c));
    TRY(P.upload(&L.rw, rw));
    TRY(P.upload(&L.cptr, cnt));
    TRY(P.upload(&L.clist, clist));
  }
  L.uploaded = true;
  out = std::move(Lp);
  return 0;
}

void drop_cycle_graph(flite_ctx *c) {
  if (c->cycle_exec) (void)hipGraphExecDestroy(c->cycle_exec);
  c->cycle_exec = nullptr;
}

int upload_level(flite_ctx *c, int level, const flite_level_view *v) {
  if (level > 1) {
    if (c->staged.size() < (size_t)level - 1 || !c->staged_ok[level - 2])
      return fail(FLITE_E_STATE, "upload the finer level first");
    if (v && v->n_fine != c->staged[level - 2].n) return fail(FLITE_E_ARG, "n_fine does not match the finer level");
  }
  Staged S;
  TRY(stage_level(c, level, v, S));
  c->staged[level - 1] = std::move(S);
  c->staged_ok[level - 1] = 1;
  c->built = false;
  drop_cycle_graph(c);
  for (auto &l : c->lev) l.reset();
  c->lev_alt.clear();
  return 0;
}

// a level's device array of nbatch x nc blocks of n values <-> host numbering
int carry_out(const Level &L, const void *d, int nc, std::vector<double> &h) {
  const size_t nblk_ = (size_t)nc * L.nbatch, cnt = nblk_ * L.n;
  std::vector<double> raw(cnt);
  if (L.fp32) {
    std::vector<float> t(cnt);
    HIPTRY(hipMemcpy(t.data(), d, cnt * 4, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < cnt; ++k) raw[k] = t[k];
  } else {
    HIPTRY(hipMemcpy(raw.data(), d, cnt * 8, hipMemcpyDeviceToHost));
  }
  h.resize(cnt);
  for (size_t b = 0; b < nblk_; ++b)
    for (int hh = 0; hh < L.n; ++hh) h[b * L.n + hh] = raw[b * L.n + L.h2d[hh]];
  return 0;
}

int carry_in(const Level &L, void *d, int nc, const std::vector<double> &h) {
  const size_t nblk_ = (size_t)nc * L.nbatch, cnt = nblk_ * L.n;
  if (h.size() != cnt) return fail(FLITE_E_STATE, "carried solution does not fit the rebuilt level");
  std::vector<double> raw(cnt);
  for (size_t b = 0; b < nblk_; ++b)
    for (int hh = 0; hh < L.n; ++hh) raw[b * L.n + L.h2d[hh]] = h[b * L.n + hh];
  if (L.fp32) {
    std::vector<float> t(raw.begin(), raw.end());
    HIPTRY(hipMemcpy(d, t.data(), cnt * 4, hipMemcpyHostToDevice));
  } else {
    HIPTRY(hipMemcpy(d, raw.data(), cnt * 8, hipMemcpyHostToDevice));
  }
  return 0;
}

// builds the device levels once every level is staged
int ensure_built(flite_ctx *c) {
  if (c->built) return 0;
  for (int l = 0; l < c->nl; ++l)
    if (!c->staged_ok[l]) return fail(FLITE_E_STATE, "not all levels uploaded");
  c->orders = nested_orders(c->staged);
  c->lev_alt.clear();
  for (int l = 0; l < c->nl; ++l)
    TRY(build_level(c, l + 1, c->staged[l], c->orders[l], c->hp.coarse_fp32 != 0, c->lev[l]));
  c->built = true;
  if (!c->carry.empty()) {  // a moved mesh: the solution arrays carry over
    auto carry = std::move(c->carry);
    c->carry.clear();
    for (int l = 0; l < c->nl; ++l) {
      Level &L = *c->lev[l];
      TRY(carry_in(L, L.u, 6, carry[l][0]));
      TRY(carry_in(L, L.uold, 6, carry[l][1]));
      TRY(carry_in(L, L.uold2, 6, carry[l][2]));
      TRY(carry_in(L, L.mu, 1, carry[l][3]));
    }
    // the step's preamble (MG:586-601) follows its readComputationData
    // (MG:545): coarse u and uold restricted with the new transfer weights
    c->preamble_due = true;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// operations
// ---------------------------------------------------------------------------
int op_prep(flite_ctx *c, Level &L, bool compute_p) {
  DevParams P = c->P;
  Timed tm(c, "prep", L.gnum);
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    if (compute_p) {
      hipLaunchKernelGGL(k_prep<R>, dim3(nblk(L.n) * L.nbatch), dim3(NTHR), 0, c->stream, V, P);
    } else {
      hipLaunchKernelGGL(k_pack<R>, dim3(nblk(L.n) * L.nbatch), dim3(NTHR), 0, c->stream, V, P);
    }
    return check_launch();
  });
}

int op_grad(flite_ctx *c, Level &L) {
  if (!(L.has_coords && c->hp.hllc_flux_order > 1 && c->hp.flux_type == 4)) return 0;
  DevParams P = c->P;
  Timed tm(c, "grad", L.gnum);
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    if (L.lanes == 1 && !c->grad_direct) {
      hipLaunchKernelGGL(k_grad_minmax_st<R>, dim3(nblk(L.n) * L.nbatch), dim3(NTHR), 0, c->stream, V, P);
      return check_launch();
    }
    return with_lanes(L.lanes, [&](auto sc) {
      constexpr int SUB = decltype(sc)::value;
      hipLaunchKernelGGL((k_grad_minmax<R, SUB>), dim3(nblk((long long)L.n * SUB) * L.nbatch), dim3(NTHR), 0, c->stream, V, P);
      return check_launch();
    });
  });
}

int op_flux(flite_ctx *c, Level &L) {
  DevParams P = c->P;
  bool o2 = L.has_coords && c->hp.hllc_flux_order > 1;
  if (L.gnum <= (int)c->evals.size()) c->evals[L.gnum - 1] += 1;
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    if (L.ne > 0 && P.ftype != 4) {
      // central / LLF / Roe: one flux per side into fbuf
      Timed tm(c, "flux", L.gnum);
      dim3 g(nblk(L.ne) * L.nbatch), b(NTHR);
      L.flux_partials = false;
#define LAUNCH_SIMPLE(FT)                                                                        \
  if (P.sa) hipLaunchKernelGGL((k_flux_simple<R, FT, true>), g, b, 0, c->stream, V, P);          \
  else hipLaunchKernelGGL((k_flux_simple<R, FT, false>), g, b, 0, c->stream, V, P);
      if (P.ftype == 1) {
        LAUNCH_SIMPLE(1);
      } else if (P.ftype == 2) {
        LAUNCH_SIMPLE(2);
      } else {
        LAUNCH_SIMPLE(3);
      }
#undef LAUNCH_SIMPLE
    } else if (L.ne > 0) {
      Timed tm(c, "flux", L.gnum);
      dim3 g(nblk(L.ne) * L.nbatch), b(NTHR);
      bool n4 = P.nsig == 4;  // LED exponent: 4 inviscid, 2 viscous
      L.flux_partials = !c->flux_per_side;
      if (L.flux_partials) {
        auto blk = [&](auto shape) {
          constexpr int SH = decltype(shape)::value;
          dim3 gb((unsigned)L.sb[SH].nfb * L.nbatch), bb(FB_SHAPE_THR[SH]);
          if (o2 && P.sa) {
            if (n4) hipLaunchKernelGGL((k_flux_blk<R, 2, true, 4, SH>), gb, bb, 0, c->stream, V, P);
            else hipLaunchKernelGGL((k_flux_blk<R, 2, true, 2, SH>), gb, bb, 0, c->stream, V, P);
          } else if (o2) {
            if (n4) hipLaunchKernelGGL((k_flux_blk<R, 2, false, 4, SH>), gb, bb, 0, c->stream, V, P);
            else hipLaunchKernelGGL((k_flux_blk<R, 2, false, 2, SH>), gb, bb, 0, c->stream, V, P);
          } else if (P.sa) {
            hipLaunchKernelGGL((k_flux_blk<R, 1, true, 2, SH>), gb, bb, 0, c->stream, V, P);
          } else {
            hipLaunchKernelGGL((k_flux_blk<R, 1, false, 2, SH>), gb, bb, 0, c->stream, V, P);
          }
        };
        if (L.fb_shape == 0) blk(std::integral_constant<int, 0>{});
        else blk(std::integral_constant<int, 1>{});
      } else if (o2 && P.sa) {
        if (n4) hipLaunchKernelGGL((k_flux<R, 2, true, 4>), g, b, 0, c->stream, V, P);
        else hipLaunchKernelGGL((k_flux<R, 2, true, 2>), g, b, 0, c->stream, V, P);
      } else if (o2) {
        if (n4) hipLaunchKernelGGL((k_flux<R, 2, false, 4>), g, b, 0, c->stream, V, P);
        else hipLaunchKernelGGL((k_flux<R, 2, false, 2>), g, b, 0, c->stream, V, P);
      } else if (P.sa) {
        hipLaunchKernelGGL((k_flux<R, 1, true, 2>), g, b, 0, c->stream, V, P);
      } else {
        hipLaunchKernelGGL((k_flux<R, 1, false, 2>), g, b, 0, c->stream, V, P);
      }
    }
    if (L.nbs > 0) {
      Timed tm(c, "bflux", L.gnum);
      hipLaunchKernelGGL(k_bflux<R>, dim3(nblk(L.nbs) * L.nbatch), dim3(NTHR), 0, c->stream, V, P);
    }
    return check_launch();
  });
}

int op_stage(flite_ctx *c, Level &L, int mode, const StageArgs &S0) {
  DevParams P = c->P;
  StageArgs S = S0;
  S.partials = L.flux_partials;
  Timed tm(c, "stage", L.gnum);
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    return with_lanes(L.lanes, [&](auto sc) {
      constexpr int SUB = decltype(sc)::value;
      dim3 g(nblk((long long)L.n * SUB) * L.nbatch), t(NTHR);
#define LAUNCH_STAGE(M)                                                                       \
  if (P.sa) hipLaunchKernelGGL((k_stage<R, M, true, SUB>), g, t, 0, c->stream, V, P, S);      \
  else hipLaunchKernelGGL((k_stage<R, M, false, SUB>), g, t, 0, c->stream, V, P, S);
      switch (mode) {
        case MODE_RHS: LAUNCH_STAGE(MODE_RHS); break;
        case MODE_STAGE: LAUNCH_STAGE(MODE_STAGE); break;
        case MODE_COARSE: LAUNCH_STAGE(MODE_COARSE); break;
        default: LAUNCH_STAGE(MODE_FINE); break;
      }
#undef LAUNCH_STAGE
      return check_launch();
    });
  });
}

int op_timesteps(flite_ctx *c, Level &L) {
  DevParams P = c->P;
  Timed tm(c, "timesteps", L.gnum);
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    launch_gather<TimeStepPart<R>, NoPart<R>, NoPart<R>>(c, L, V, P);
    return check_launch();
  });
}

int op_zero_diss(flite_ctx *c, Level &L) {
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    size_t cnt = 5 * (size_t)L.n * L.nbatch;
    hipLaunchKernelGGL(k_fill<R>, dim3(nblk((long long)cnt)), dim3(NTHR), 0, c->stream, V.diss, R(0), cnt);
    return check_launch();
  });
}

// makeArtificialDissipation2 (DEF:11484-11993) in two steps: op_dissipation
// zeroes the array (as the routine does first), op_jst adds the JST terms
// once the local time steps exist.  The reference computes the time steps
// before the dissipation and subtracts the viscous term after it
// (DEF:1023-1047).  Here the viscous gather writes dissipation = -viscous
// (ViscFluxPart), so op_jst always runs after it and adds its node totals:
// -viscous + JST is the reference's JST - viscous, each a single addition.
int op_dissipation(flite_ctx *c, Level &L) { return op_zero_diss(c, L); }

bool jst_active(const flite_ctx *c, const Level &L) {
  const flite_params &hp = c->hp;
  if (hp.flux_type == 4 && L.gnum == 1) return false;  // DEF:11506
  return !(hp.fourth_order_dissipation_factor == 0.0 && hp.second_order_dissipation_factor == 0.0);  // DEF:11508
}

int op_jst(flite_ctx *c, Level &L) {
  if (!jst_active(c, L) || L.ne == 0) return 0;
  DevParams P = c->P;
  Timed tm(c, "dissipation", L.gnum);
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    dim3 g(nblk(L.n) * L.nbatch), b(NTHR);
    hipLaunchKernelGGL(k_jst_sums<R>, g, b, 0, c->stream, V, P);
    if (P.matdiss) hipLaunchKernelGGL((k_jst_flux<R, true>), g, b, 0, c->stream, V, P);
    else hipLaunchKernelGGL((k_jst_flux<R, false>), g, b, 0, c->stream, V, P);
    return check_launch();
  });
}

int op_viscous(flite_ctx *c, Level &L) {
  DevParams P = c->P;
  Timed tm(c, "viscous", L.gnum);
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    launch_gather<ViscNodePart<R, true>, NoPart<R>, NoPart<R>>(c, L, V, P);
    launch_gather<ViscFluxPart<R>, NoPart<R>, NoPart<R>>(c, L, V, P);
    return check_launch();
  });
}

int op_sa_timesteps(flite_ctx *c, Level &L) {
  DevParams P = c->P;
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    launch_gather<SATimeStepPart<R>, NoPart<R>, NoPart<R>>(c, L, V, P);
    return check_launch();
  });
}

int op_sa_diffusion(flite_ctx *c, Level &L) {
  DevParams P = c->P;
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    launch_gather<SADiff1Part<R>, NoPart<R>, NoPart<R>>(c, L, V, P);
    launch_gather<SADiff2Part<R>, NoPart<R>, NoPart<R>>(c, L, V, P);
    return check_launch();
  });
}

// The viscous-flow terms of one smoothing step in two walks of the
// incidence lists instead of up to six:
//   pass A: [makeTimeSteps] + makeViscosityTerm pass 1 + [makeSADiffusion pass 1]
//   pass B: makeViscosityTerm pass 2 + [makeTurbulentTimeSteps] + [makeSADiffusion pass 2]
// Pass B only needs pass A's node results (mu, tau, vorticity, divergence,
// grad nu~), which is the dependency order of the separate routines.  With
// time steps in pass A they read the previous laminar viscosity, so the new
// one reaches the gather records after the pass (k_mu_record).
int op_step_terms(flite_ctx *c, Level &L, bool timesteps, bool sa_timesteps) {
  DevParams P = c->P;
  bool sa = P.sa;
  Timed tm(c, "viscous", L.gnum);
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    if (timesteps && sa) {
      launch_gather<TimeStepPart<R>, ViscNodePart<R, false>, SADiff1Part<R>>(c, L, V, P);
    } else if (timesteps) {
      launch_gather<TimeStepPart<R>, ViscNodePart<R, false>, NoPart<R>>(c, L, V, P);
    } else if (sa) {
      launch_gather<ViscNodePart<R, true>, SADiff1Part<R>, NoPart<R>>(c, L, V, P);
    } else {
      launch_gather<ViscNodePart<R, true>, NoPart<R>, NoPart<R>>(c, L, V, P);
    }
    if (timesteps)
      hipLaunchKernelGGL(k_mu_record<R>, dim3(nblk(L.n) * L.nbatch), dim3(NTHR), 0, c->stream, V);
    if (sa && sa_timesteps) {
    launch_gather<ViscFluxPart<R>, SATimeStepPart<R>, SADiff2Part<R>>(c, L, V, P);
    } else if (sa) {
    launch_gather<ViscFluxPart<R>, SADiff2Part<R>, NoPart<R>>(c, L, V, P);
    } else {
    launch_gather<ViscFluxPart<R>, NoPart<R>, NoPart<R>>(c, L, V, P);
    }
    return check_launch();
  });
}

int op_copy(flite_ctx *c, Level &L, void *dst, const void *src, int ncomp) {
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    hipLaunchKernelGGL(k_copy_state<R>, dim3(nblk(L.n) * L.nbatch), dim3(NTHR), 0, c->stream, (R *)dst, (const R *)src,
                       L.n, ncomp, L.nbatch);
    return check_launch();
  });
}

int op_bc_state(flite_ctx *c, Level &L, void *a, int outer) {
  DevParams P = c->P;
  return with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    hipLaunchKernelGGL(k_bc_state<R>, dim3(nblk(L.n) * L.nbatch), dim3(NTHR), 0, c->stream, V, P, (R *)a, outer);
    return check_launch();
  });
}

// mapFineToCoarse of ncomp columns from array fa of F into array ka of K
int op_restrict_array(flite_ctx *c, Level &F, Level &K, const void *fa, void *ka, int ncomp) {
  auto go = [&](auto *fin, auto *kout) {
    using RF = std::remove_const_t<std::remove_pointer_t<decltype(fin)>>;
    using RC = std::remove_pointer_t<decltype(kout)>;
    constexpr int SUB = 4;  // about 8 fine members per coarse node
    long long per = ncomp == 1 ? 1 : 6;  // components allocated per node of the array
    hipLaunchKernelGGL((k_restrict<RF, RC, SUB>), dim3(nblk((long long)K.n * SUB) * K.nbatch), dim3(NTHR), 0,
                       c->stream, K.cptr, K.rw, fin, kout, K.n, F.n, ncomp, K.nbatch, per * F.n, per * K.n);
    return check_launch();
  };
  if (F.fp32 && K.fp32) return go((const float *)fa, (float *)ka);
  if (!F.fp32 && K.fp32) return go((const double *)fa, (float *)ka);
  if (F.fp32 && !K.fp32) return go((const float *)fa, (double *)ka);
  return go((const double *)fa, (double *)ka);
}

int do_single_iteration(flite_ctx *c, Level &L, bool calc_diss) {
  const flite_params &hp = c->hp;
  if (calc_diss) {
    TRY(op_prep(c, L, true));
    TRY(op_dissipation(c, L));
    if (hp.reynolds_number > 0.0) {
      // makeTimeSteps, makeViscosityTerm, makeTurbulentTimeSteps, makeSADiffusion
      TRY(op_step_terms(c, L, true, true));
    } else {
      TRY(op_timesteps(c, L));
    }
    TRY(op_jst(c, L));
  }
  for (int k = 1; k <= hp.rk_stages; ++k) {
    // stage k > 1 starts from the pressure and records the previous stage
    // kernel wrote with the new state; stage 1 after the frozen terms too
    if (k == 1 && !calc_diss) TRY(op_prep(c, L, true));
    TRY(op_grad(c, L));
    TRY(op_flux(c, L));
    StageArgs S;
    S.alpha = hp.rk_coefficients[k - 1];
    S.dtj = hp.cfl_number * S.alpha;
    S.dtj_t = hp.turbulent_cfl_number * S.alpha;
    S.calc_turb = 1;
    S.fas = L.gnum > 1;
    S.has_src = L.gnum > 1;
    S.next_record = k < hp.rk_stages;
    TRY(op_stage(c, L, MODE_STAGE, S));
  }
  return 0;
}

int do_time_iterations(flite_ctx *c, int level, int steps, bool calc_diss) {
  Level &L = *c->lev[level - 1];
  for (int s = 0; s < steps; ++s) {
    TRY(op_copy(c, L, L.uprev, L.u, c->P.sa ? 6 : 5));
    TRY(do_single_iteration(c, L, calc_diss));
  }
  return 0;
}

int do_restrict(flite_ctx *c, int fine_level) {
  Level &F = *c->lev[fine_level - 1];
  Level &K = *c->lev[fine_level];
  const flite_params &hp = c->hp;
  bool visc = hp.reynolds_number > 0.0;
  // coarse part (DEF:12158-12234)
  TRY(op_restrict_array(c, F, K, F.u, K.u, 6));
  TRY(op_restrict_array(c, F, K, F.mu, K.mu, 1));
  TRY(op_prep(c, K, true));
  TRY(op_bc_state(c, K, K.u, 1));
  TRY(op_copy(c, K, K.uprev, K.u, 6));
  TRY(op_prep(c, K, true));
  TRY(op_dissipation(c, K));
  if (visc) {
    TRY(op_step_terms(c, K, true, true));
  } else {
    TRY(op_timesteps(c, K));
  }
  TRY(op_jst(c, K));
  TRY(op_flux(c, K));
  StageArgs S;
  S.alpha = 1.0;
  S.dtj = hp.cfl_number;
  S.dtj_t = hp.turbulent_cfl_number;
  S.calc_turb = 1;
  S.fas = 0;
  S.has_src = 1;
  TRY(op_stage(c, K, MODE_COARSE, S));
  // fine part (DEF:12236-12361)
  TRY(op_prep(c, F, true));
  TRY(op_dissipation(c, F));
  if (visc) TRY(op_step_terms(c, F, false, false));
  TRY(op_jst(c, F));  // with the time steps of the last smoothing step
  TRY(op_grad(c, F));
  TRY(op_flux(c, F));
  S.has_src = F.gnum > 1;
  TRY(op_stage(c, F, MODE_FINE, S));
  // coarse forcing (DEF:12362-12428)
  TRY(op_restrict_array(c, F, K, F.rhs, K.rhs, 6));
  DevParams P = c->P;
  TRY(with_level(K, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    hipLaunchKernelGGL(k_restrict_finish<R>, dim3(nblk(K.n) * K.nbatch), dim3(NTHR), 0, c->stream, V, P);
    return check_launch();
  }));
  return 0;
}

int do_prolongate(flite_ctx *c, int coarse_level) {
  Level &K = *c->lev[coarse_level - 1];
  Level &F = *c->lev[coarse_level - 2];
  DevParams P = c->P;
  TRY(op_restrict_array(c, F, K, F.u, K.uprev, 6));
  TRY(op_bc_state(c, K, K.uprev, 1));
  TRY(with_level(K, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    hipLaunchKernelGGL(k_coarse_correction<R>, dim3(nblk(K.n) * K.nbatch), dim3(NTHR), 0, c->stream, V, P);
    return check_launch();
  }));
  int ncomp = (F.gnum < c->hp.number_of_turbulence_grids && P.sa) ? 6 : 5;
  auto go = [&](auto FV, auto *crhs) {
    using RF = std::remove_pointer_t<decltype(FV.u)>;
    using RC = std::remove_pointer_t<decltype(crhs)>;
    hipLaunchKernelGGL((k_prolong_fine<RF, RC>), dim3(nblk(F.n) * F.nbatch), dim3(NTHR), 0, c->stream, FV, P, K.f2c,
                       (const RC *)crhs, K.n, ncomp);
    return check_launch();
  };
  if (F.fp32) {
    auto FV = F.view<float>();
    return K.fp32 ? go(FV, (float *)K.rhs) : go(FV, (double *)K.rhs);
  }
  auto FV = F.view<double>();
  return K.fp32 ? go(FV, (float *)K.rhs) : go(FV, (double *)K.rhs);
}

int do_cycle(flite_ctx *c, int cycle) {
  const flite_params &hp = c->hp;
  int ng = c->nl, nrs = hp.number_of_relaxation_steps;
  if (cycle > hp.number_of_turbulence_steps && c->P.sa)
    return fail(FLITE_E_UNSUPPORTED, "frozen turbulence (cycle > numberOfTurbulenceSteps) is not implemented");
  if (hp.multigrid_scheme <= 0) {
    if (ng > 1) return fail(FLITE_E_UNSUPPORTED, "multigridScheme <= 0 with more than one grid is not implemented");
    for (int i = 1; i < ng; ++i) {
      Level &F = *c->lev[i - 1];
      TRY(op_copy(c, F, F.uprev, F.u, 6));
      TRY(op_restrict_array(c, F, *c->lev[i], F.u, c->lev[i]->u, 6));
    }
    Level &Kc = *c->lev[ng - 1];
    TRY(with_level(Kc, [&](auto V) {
      using R = std::remove_pointer_t<decltype(V.u)>;
      size_t cnt = 6 * (size_t)Kc.n * Kc.nbatch;
      hipLaunchKernelGGL(k_fill<R>, dim3(nblk((long long)cnt)), dim3(NTHR), 0, c->stream, V.src, R(0), cnt);
      return check_launch();
    }));
    TRY(do_time_iterations(c, ng, nrs, true));
  } else if (hp.multigrid_scheme == 1) {
    for (int i = 1; i < ng; ++i) {
      TRY(do_time_iterations(c, i, nrs, i == 1));
      TRY(do_restrict(c, i));
      Level &F = *c->lev[i - 1];
      TRY(op_restrict_array(c, F, *c->lev[i], F.u, c->lev[i]->u, 6));
    }
    TRY(do_time_iterations(c, ng, nrs, ng == 1));
    for (int i = ng; i >= 2; --i) {
      TRY(do_prolongate(c, i));
      TRY(do_time_iterations(c, i - 1, nrs, (i - 1) == 1));
    }
  } else if (hp.multigrid_scheme == 3) {
    int sweeps = ng == 1 ? 1 : ng - 1, finest = 1;
    for (int s = 1; s <= sweeps; ++s) {
      for (int j = finest; j <= ng - 1; ++j) {
        TRY(do_time_iterations(c, j, nrs, j == finest));
        TRY(do_restrict(c, j));
      }
      TRY(do_time_iterations(c, ng, nrs, ng == 1));
      finest = finest == 1 ? (ng >= 2 ? ng - 1 : ng) : finest - 1;
      for (int j = ng; j >= finest + 1; --j) TRY(do_prolongate(c, j));
    }
  } else {
    return fail(FLITE_E_UNSUPPORTED, "multigrid scheme not implemented");
  }
  return 0;
}

// multiGridCycles preamble (MG:586-602): laminar viscosity, u and (for
// unsteady runs) uold restricted to every coarse level
int preamble(flite_ctx *c) {
  double mu0 = c->hp.reynolds_number > 0.0 ? 1.0 / c->hp.reynolds_number : 0.0;
  for (int l = 0; l < c->nl; ++l) {
    Level &L = *c->lev[l];
    TRY(with_level(L, [&](auto V) {
      using R = std::remove_pointer_t<decltype(V.u)>;
      size_t cnt = (size_t)L.n * L.nbatch;
      hipLaunchKernelGGL(k_fill<R>, dim3(nblk((long long)cnt)), dim3(NTHR), 0, c->stream, V.mu, R(mu0), cnt);
      return check_launch();
    }));
    if (l > 0) {
      TRY(op_restrict_array(c, *c->lev[l - 1], L, c->lev[l - 1]->u, L.u, 6));
      if (c->hp.flow_type != 0) TRY(op_restrict_array(c, *c->lev[l - 1], L, c->lev[l - 1]->uold, L.uold, 6));
    }
  }
  return sync(c);
}

// uold2 <- src_old, uold <- src on level 1 (unsteady runs)
int set_old_states(flite_ctx *c, const void *old_src, const void *src) {
  if (c->hp.flow_type == 0) return 0;
  Level &L = *c->lev[0];
  TRY(op_copy(c, L, L.uold2, old_src, 6));
  return op_copy(c, L, L.uold, src, 6);
}

int check_ctx(flite_ctx *c) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  TRY(ensure_built(c));
  if (c->preamble_due) {
    c->preamble_due = false;
    TRY(preamble(c));
  }
  return 0;
}

int check_level(flite_ctx *c, int level) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  if (level < 1 || level > c->nl) return fail(FLITE_E_ARG, "level out of range");
  return check_ctx(c);
}

void *field_ptr(Level &L, int field, int *ncomp, size_t *stride_nodes) {
  *stride_nodes = (size_t)L.n;
  switch (field) {
    case FLITE_FIELD_U: *ncomp = 6; return L.u;
    case FLITE_FIELD_UPREV: *ncomp = 6; return L.uprev;
    case FLITE_FIELD_RHS: *ncomp = 6; return L.rhs;
    case FLITE_FIELD_DISSIPATION: *ncomp = 5; return L.diss;
    case FLITE_FIELD_PRESSURE: *ncomp = 1; return L.p;
    case FLITE_FIELD_DT: *ncomp = 1; return L.dt;
    case FLITE_FIELD_DT_TURB: *ncomp = 1; return L.dtt;
    case FLITE_FIELD_LAMINAR_VISCOSITY: *ncomp = 1; return L.mu;
    case FLITE_FIELD_VORTICITY: *ncomp = 1; return L.vort;
    case FLITE_FIELD_DIVERGENCE: *ncomp = 1; return L.divg;
    case FLITE_FIELD_SOURCE: *ncomp = 6; return L.src;
    case FLITE_FIELD_TURB_DIFFUSION: *ncomp = 1; return L.tdiff;
    case FLITE_FIELD_UOLD: *ncomp = 6; return L.uold;
    case FLITE_FIELD_UOLD2: *ncomp = 6; return L.uold2;
    default: return nullptr;
  }
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char *flite_last_error(void) { return g_err.c_str(); }

int flite_field_ncomp(int field) {
  static const int nc[FLITE_FIELD_COUNT] = {6, 6, 6, 5, 1, 1, 1, 1, 1, 1, 6, 1, 25, 3, 6, 6};
  return field >= 0 && field < FLITE_FIELD_COUNT ? nc[field] : -1;
}

int flite_ctx_create(const flite_params *p, int n_levels, int n_batch, flite_ctx **out) {
  if (!p || !out) return fail(FLITE_E_ARG, "null argument");
  *out = nullptr;
  if (n_levels < 1 || n_levels > 16) return fail(FLITE_E_ARG, "n_levels must be 1..16");
  if (n_batch < 1) return fail(FLITE_E_ARG, "n_batch must be >= 1");
  // a batch varies only the freestream: every other setting must match
  for (int b = 1; b < n_batch; ++b) {
    const flite_params &q = p[b];
    bool same = q.gamma == p->gamma && q.reynolds_number == p->reynolds_number &&
                q.prandtl_number == p->prandtl_number && q.turbulent_prandtl_number == p->turbulent_prandtl_number &&
                q.inflow_temperature == p->inflow_temperature && q.walls_are_isentropic == p->walls_are_isentropic &&
                q.turbulence_model == p->turbulence_model && q.flux_type == p->flux_type &&
                q.hllc_flux_order == p->hllc_flux_order && q.automatic_clipping == p->automatic_clipping &&
                q.gradient_method == p->gradient_method && q.hllc_clipping == p->hllc_clipping &&
                q.boundary_term == p->boundary_term && q.cfl_number == p->cfl_number &&
                q.turbulent_cfl_number == p->turbulent_cfl_number && q.multigrid_scheme == p->multigrid_scheme &&
                q.number_of_relaxation_steps == p->number_of_relaxation_steps &&
                q.number_of_turbulence_grids == p->number_of_turbulence_grids &&
                q.number_of_turbulence_steps == p->number_of_turbulence_steps &&
                q.max_increment_factor == p->max_increment_factor && q.minimum_density == p->minimum_density &&
                q.max_turbulence_value == p->max_turbulence_value && q.turb_k == p->turb_k &&
                q.prolongation_relaxation == p->prolongation_relaxation &&
                q.turb_prolongation_relaxation == p->turb_prolongation_relaxation &&
                q.multigrid_bc_relaxation == p->multigrid_bc_relaxation && q.coarse_fp32 == p->coarse_fp32 &&
                q.second_order_dissipation_factor == p->second_order_dissipation_factor &&
                q.fourth_order_dissipation_factor == p->fourth_order_dissipation_factor &&
                q.coarse_grid_dissipation_factor == p->coarse_grid_dissipation_factor &&
                q.dissipation_scheme == p->dissipation_scheme &&
                q.coarse_grid_dissipation_scheme == p->coarse_grid_dissipation_scheme &&
                q.use_matrix_dissipation == p->use_matrix_dissipation &&
                q.hartens_correction_factor == p->hartens_correction_factor &&
                q.flow_type == p->flow_type && q.time_scheme == p->time_scheme &&
                q.add_time_step == p->add_time_step && q.physical_time_step == p->physical_time_step;
    same = same && q.rk_stages == p->rk_stages;
    for (int k = 0; k < 5; ++k) same = same && q.rk_coefficients[k] == p->rk_coefficients[k];
    for (int k = 0; k < 6; ++k) same = same && q.inflow_field[1][k] == p->inflow_field[1][k];
    if (!same)
      return fail(FLITE_E_ARG, "batched conditions may differ only in Mach number, alpha, beta and the inflow state");
  }
  if (p->flux_type < 1 || p->flux_type > 4) return fail(FLITE_E_UNSUPPORTED, "fluxType must be 1-4");
  if (p->rk_stages < 1 || p->rk_stages > 5) return fail(FLITE_E_ARG, "rk_stages must be 1..5");
  if (p->flow_type < 0 || p->flow_type > 2)
    return fail(FLITE_E_UNSUPPORTED, "flowType must be 0 (steady), 1 (fixed mesh) or 2 (moving mesh)");
  if (p->flow_type != 0 && !(p->physical_time_step > 0.0)) return fail(FLITE_E_ARG, "physical_time_step must be > 0");
  if (p->flux_type != 4 && (p->dissipation_scheme != 2 || p->coarse_grid_dissipation_scheme != 2))
    return fail(FLITE_E_UNSUPPORTED, "only dissipation scheme 2 (JST) is implemented");
  if (p->turbulence_model != 0 && p->turbulence_model != 1)
    return fail(FLITE_E_UNSUPPORTED, "only turbulenceModel 0/1 are implemented");
  if (p->hllc_flux_order > 1 && p->gradient_method != 1)
    return fail(FLITE_E_UNSUPPORTED, "only Green-Gauss gradients are implemented");
  if (p->turbulence_model == 1 && !(p->reynolds_number > 0.0))
    return fail(FLITE_E_ARG, "Spalart-Allmaras needs ReynoldsNumber > 0");
  int ndev = 0;
  HIPTRY(hipGetDeviceCount(&ndev));
  if (ndev < 1) return fail(FLITE_E_HIP, "no HIP device");
  int dev = 0;
  HIPTRY(hipGetDevice(&dev));
  TRY(scratch_setup(dev));  // this device's copy of the monitor scratch, once per device
  auto c = std::make_unique<flite_ctx>();
  c->device = dev;
  c->hp = *p;
  c->P = make_dev_params(*p);
  c->P.cond = nullptr;
  c->hp_b.assign(p, p + n_batch);
  if (n_batch > 1) {
    std::vector<double> cond(8 * (size_t)n_batch, 0.0);
    for (int b = 0; b < n_batch; ++b) {
      cond[8 * (size_t)b] = p[b].mach_number;
      for (int k = 0; k < 6; ++k) cond[8 * (size_t)b + 1 + k] = p[b].inflow_field[0][k];
    }
    HIPTRY(hipMalloc(&c->cond_dev, cond.size() * sizeof(double)));
    HIPTRY(hipMemcpy(c->cond_dev, cond.data(), cond.size() * sizeof(double), hipMemcpyHostToDevice));
    c->P.cond = c->cond_dev;
  }
  if (const char *e = std::getenv("FLITE_GATHER_LANES")) {
    int v = std::atoi(e);
    if (v != 1 && v != 4 && v != 16) {
      return fail(FLITE_E_ARG, "FLITE_GATHER_LANES must be 1, 4 or 16");
    }
    c->lanes_override = v;
  }
  if (const char *e = std::getenv("FLITE_SIDE_BLOCK")) {
    int v = std::atoi(e);
    for (int k = 0; k < FB_NSHAPES; ++k)
      if (v == FB_SHAPE_THR[k]) c->shape_override = k;
    if (c->shape_override < 0) return fail(FLITE_E_ARG, "FLITE_SIDE_BLOCK must be 256 or 512");
  }
  if (const char *e = std::getenv("FLITE_FLUX_KERNEL")) {
    std::string v(e);
    if (v != "side" && v != "block") return fail(FLITE_E_ARG, "FLITE_FLUX_KERNEL must be side or block");
    c->flux_per_side = v == "side";
  }
  if (const char *e = std::getenv("FLITE_CYCLE_GRAPH")) {
    std::string v(e);
    if (v != "0" && v != "1") return fail(FLITE_E_ARG, "FLITE_CYCLE_GRAPH must be 0 or 1");
    c->graph_off = v == "0";
  }
  if (const char *e = std::getenv("FLITE_GRAD_STORE")) {
    std::string v(e);
    if (v != "direct" && v != "lds") return fail(FLITE_E_ARG, "FLITE_GRAD_STORE must be direct or lds");
    c->grad_direct = v == "direct";
  }
  if (const char *e = std::getenv("FLITE_GATHER_KERNEL")) {
    std::string v(e);
    if (v != "node" && v != "block") return fail(FLITE_E_ARG, "FLITE_GATHER_KERNEL must be node or block");
    c->gather_per_node = v == "node";
  }
  c->nl = n_levels;
  c->nbatch = n_batch;
  c->lev.resize(n_levels);
  c->staged.resize(n_levels);
  c->staged_ok.assign(n_levels, 0);
  HIPTRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  *out = c.release();
  return 0;
}

void flite_ctx_destroy(flite_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_cycle_graph(c);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (auto &pr : c->ev_used) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  c->lev.clear();
  c->lev_alt.clear();
  if (c->cond_dev) (void)hipFree(c->cond_dev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int flite_level_upload(flite_ctx *c, int level, const flite_level_view *v) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  if (level < 1 || level > c->nl) return fail(FLITE_E_ARG, "level out of range");
  TRY(sync(c));
  return upload_level(c, level, v);
}

int flite_level_geometry(flite_ctx *c, int level, const flite_level_view *v) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  if (level < 1 || level > c->nl) return fail(FLITE_E_ARG, "level out of range");
  if (c->hp.flow_type != 2) return fail(FLITE_E_STATE, "a moving mesh needs flow_type 2");
  if (!c->staged_ok[level - 1]) return fail(FLITE_E_STATE, "level not uploaded");
  const Staged &O = c->staged[level - 1];
  if (!v) return fail(FLITE_E_ARG, "null level view");
  if (v->n_nodes != O.n || v->n_sides != O.ne || v->n_bsides != O.nbs || v->n_bnodes != O.nb)
    return fail(FLITE_E_ARG, "the moved mesh must keep the level's node, side and boundary counts");
  if (c->built && c->carry.empty()) {
    TRY(sync(c));
    c->carry.assign(c->nl, std::vector<std::vector<double>>(4));
    for (int l = 0; l < c->nl; ++l) {
      const Level &L = *c->lev[l];
      TRY(carry_out(L, L.u, 6, c->carry[l][0]));
      TRY(carry_out(L, L.uold, 6, c->carry[l][1]));
      TRY(carry_out(L, L.uold2, 6, c->carry[l][2]));
      TRY(carry_out(L, L.mu, 1, c->carry[l][3]));
    }
  }
  if (c->built) TRY(sync(c));
  return upload_level(c, level, v);
}

int flite_field_get(flite_ctx *c, int level, int field, double *out, int ld) {
  TRY(check_level(c, level));
  Level &L = *c->lev[level - 1];
  if (!out || ld < L.n) return fail(FLITE_E_ARG, "bad output buffer");
  TRY(sync(c));
  const size_t es = L.fp32 ? 4 : 8;
  auto to_host = [&](const void *dptr, size_t count, std::vector<double> &h) -> int {
    h.resize(count);
    if (L.fp32) {
      std::vector<float> t(count);
      HIPTRY(hipMemcpy(t.data(), dptr, count * es, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < count; ++k) h[k] = t[k];
    } else {
      HIPTRY(hipMemcpy(h.data(), dptr, count * es, hipMemcpyDeviceToHost));
    }
    return 0;
  };
  std::vector<double> h;
  const size_t sb = (size_t)c->sel;  // selected condition
  if (field == FLITE_FIELD_GRADIENT) {
    if (!(L.has_coords && c->hp.hllc_flux_order > 1)) return fail(FLITE_E_STATE, "no gradient on this level");
    TRY(to_host((char *)L.grec + sb * (size_t)L.gstride * es, GREC * (size_t)L.n, h));
    for (int hh = 0; hh < L.n; ++hh) {
      const double *g = &h[GREC * (size_t)L.h2d[hh]];
      for (int k = 0; k < 15; ++k) out[(size_t)k * ld + hh] = g[k];
      for (int k = 0; k < 5; ++k) {  // limits from their stored reciprocals
        out[(size_t)(15 + k) * ld + hh] = g[15 + k] != 0.0 ? 1.0 / g[15 + k] : 0.0;
        out[(size_t)(20 + k) * ld + hh] = g[20 + k] != 0.0 ? 1.0 / g[20 + k] : 0.0;
      }
    }
    return 0;
  }
  if (field == FLITE_FIELD_WALL_STRESS) {
    TRY(to_host((char *)L.ws + sb * 3 * (size_t)L.n * es, 3 * (size_t)L.n, h));
    for (int hh = 0; hh < L.n; ++hh)
      for (int k = 0; k < 3; ++k) out[(size_t)k * ld + hh] = h[3 * (size_t)L.h2d[hh] + k];
    return 0;
  }
  int nc;
  size_t stride;
  if ((field == FLITE_FIELD_UOLD || field == FLITE_FIELD_UOLD2) && c->hp.flow_type == 0)
    return fail(FLITE_E_STATE, "no solutions of earlier physical times in a steady run");
  void *p = field_ptr(L, field, &nc, &stride);
  if (!p) return fail(FLITE_E_ARG, "bad field");
  p = (char *)p + sb * nc * stride * es;
  TRY(to_host(p, nc * stride, h));
  for (int k = 0; k < nc; ++k)
    for (int hh = 0; hh < L.n; ++hh) out[(size_t)k * ld + hh] = h[(size_t)k * stride + L.h2d[hh]];
  return 0;
}

int flite_field_set(flite_ctx *c, int level, int field, const double *in, int ld) {
  TRY(check_level(c, level));
  Level &L = *c->lev[level - 1];
  if (!in || ld < L.n) return fail(FLITE_E_ARG, "bad input buffer");
  int nc;
  size_t stride;
  if ((field == FLITE_FIELD_UOLD || field == FLITE_FIELD_UOLD2) && c->hp.flow_type == 0)
    return fail(FLITE_E_STATE, "no solutions of earlier physical times in a steady run");
  void *p = field_ptr(L, field, &nc, &stride);
  if (!p) return fail(FLITE_E_ARG, "field cannot be set");
  p = (char *)p + (size_t)c->sel * nc * stride * (L.fp32 ? 4 : 8);
  TRY(sync(c));
  std::vector<double> h(nc * stride);
  for (int k = 0; k < nc; ++k)
    for (int hh = 0; hh < L.n; ++hh) h[(size_t)k * stride + L.h2d[hh]] = in[(size_t)k * ld + hh];
  if (L.fp32) {
    std::vector<float> t(h.begin(), h.end());
    HIPTRY(hipMemcpy(p, t.data(), t.size() * 4, hipMemcpyHostToDevice));
  } else {
    HIPTRY(hipMemcpy(p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  }
  return 0;
}

int flite_state_set(flite_ctx *c, int level, const double *u, int ld, int ncomp, int batch) {
  TRY(check_level(c, level));
  if (batch < 0 || batch >= c->nbatch) return fail(FLITE_E_ARG, "batch index out of range");
  struct Sel {
    flite_ctx *c;
    int old;
    ~Sel() { c->sel = old; }
  } keep{c, c->sel};
  c->sel = batch;
  Level &L = *c->lev[level - 1];
  if (ncomp < 5 || ncomp > 6) return fail(FLITE_E_ARG, "ncomp must be 5 or 6");
  std::vector<double> full(6 * (size_t)L.n, 0.0);
  for (int k = 0; k < ncomp; ++k)
    for (int i = 0; i < L.n; ++i) full[(size_t)k * L.n + i] = u[(size_t)k * ld + i];
  return flite_field_set(c, level, FLITE_FIELD_U, full.data(), L.n);
}

int flite_state_get(flite_ctx *c, int level, double *u, int ld, int ncomp, int batch) {
  TRY(check_level(c, level));
  if (batch < 0 || batch >= c->nbatch) return fail(FLITE_E_ARG, "batch index out of range");
  struct Sel {
    flite_ctx *c;
    int old;
    ~Sel() { c->sel = old; }
  } keep{c, c->sel};
  c->sel = batch;
  Level &L = *c->lev[level - 1];
  if (ncomp < 1 || ncomp > 6) return fail(FLITE_E_ARG, "ncomp must be 1..6");
  std::vector<double> full(6 * (size_t)L.n);
  TRY(flite_field_get(c, level, FLITE_FIELD_U, full.data(), L.n));
  for (int k = 0; k < ncomp; ++k)
    for (int i = 0; i < L.n; ++i) u[(size_t)k * ld + i] = full[(size_t)k * L.n + i];
  return 0;
}

int flite_init_from_state(flite_ctx *c) {
  TRY(check_ctx(c));
  // a restart (DEF:3866-3873): uold and uold2 take the given state
  TRY(set_old_states(c, c->lev[0]->u, c->lev[0]->u));
  return preamble(c);
}

int flite_advance_physical_time(flite_ctx *c) {
  TRY(check_ctx(c));
  if (c->hp.flow_type == 0) return fail(FLITE_E_STATE, "steady run (flowType 0): no physical time");
  // a captured cycle reads uold and uold2 where they are, so it stays valid
  TRY(set_old_states(c, c->lev[0]->uold, c->lev[0]->u));
  // the next step's preamble (MG:586-601 sit inside the physical-step loop):
  // laminar viscosity reset, u and uold restricted to every coarse level
  return preamble(c);
}

int flite_init_freestream(flite_ctx *c) {
  TRY(check_ctx(c));
  Level &L = *c->lev[0];
  std::vector<double> u(6 * (size_t)L.n, 0.0);
  for (int b = 0; b < c->nbatch; ++b) {
    for (int k = 0; k < 5; ++k)
      for (int i = 0; i < L.n; ++i) u[(size_t)k * L.n + i] = c->hp_b[b].inflow_field[0][k];
    if (c->P.sa)
      for (int i = 0; i < L.n; ++i) u[5 * (size_t)L.n + i] = 0.1;
    TRY(flite_state_set(c, 1, u.data(), L.n, 6, b));
  }
  // uold and uold2 take the freestream before the boundary conditions act
  // on u (DEF:3980-3994)
  TRY(set_old_states(c, L.u, L.u));
  TRY(op_prep(c, L, true));
  TRY(op_bc_state(c, L, L.u, 0));
  return preamble(c);
}

int flite_do_time_iterations(flite_ctx *c, int level, int steps, int calc_diss, int calc_turb, int iteration) {
  TRY(check_level(c, level));
  (void)iteration;
  if (c->P.sa && !calc_turb) return fail(FLITE_E_UNSUPPORTED, "calculateTurbulence=.false. is not implemented");
  TRY(do_time_iterations(c, level, steps, calc_diss != 0));
  return sync(c);
}

int flite_restrict(flite_ctx *c, int fine_level, int calc_turb) {
  TRY(check_ctx(c));
  if (fine_level < 1 || fine_level >= c->nl) return fail(FLITE_E_ARG, "fine_level out of range");
  if (c->P.sa && !calc_turb) return fail(FLITE_E_UNSUPPORTED, "calculateTurbulence=.false. is not implemented");
  TRY(do_restrict(c, fine_level));
  return sync(c);
}

int flite_prolongate(flite_ctx *c, int coarse_level) {
  TRY(check_ctx(c));
  if (coarse_level < 2 || coarse_level > c->nl) return fail(FLITE_E_ARG, "coarse_level out of range");
  TRY(do_prolongate(c, coarse_level));
  return sync(c);
}

int flite_map_u_fine_to_coarse(flite_ctx *c, int fine_level) {
  TRY(check_ctx(c));
  if (fine_level < 1 || fine_level >= c->nl) return fail(FLITE_E_ARG, "fine_level out of range");
  Level &F = *c->lev[fine_level - 1];
  TRY(op_restrict_array(c, F, *c->lev[fine_level], F.u, c->lev[fine_level]->u, 6));
  return sync(c);
}

int flite_cycle(flite_ctx *c, int cycle_number) {
  TRY(check_ctx(c));
  if (cycle_number > c->hp.number_of_turbulence_steps && c->P.sa)
    return fail(FLITE_E_UNSUPPORTED, "frozen turbulence (cycle > numberOfTurbulenceSteps) is not implemented");
  // kernel timing records events around single launches: launched one by one
  // (events recorded while capturing cannot be timed on this runtime)
  if (c->graph_off || !c->timed.empty()) {
    c->evals.assign(c->nl, 0);
    TRY(do_cycle(c, cycle_number));
    return sync(c);
  }
  if (!c->cycle_exec) {
    c->evals.assign(c->nl, 0);
    HIPTRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = do_cycle(c, cycle_number);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != 0 || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc != 0 ? rc : fail(FLITE_E_HIP, std::string("cycle capture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&c->cycle_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      c->cycle_exec = nullptr;
      return fail(FLITE_E_HIP, std::string("cycle graph: ") + hipGetErrorString(e));
    }
    c->cycle_evals = c->evals;
  }
  c->evals = c->cycle_evals;
  HIPTRY(hipGraphLaunch(c->cycle_exec, c->stream));
  return sync(c);
}

int flite_apply(flite_ctx *c, int level, int op) {
  TRY(check_level(c, level));
  Level &L = *c->lev[level - 1];
  switch (op) {
    case FLITE_OP_PRESSURE:
    case FLITE_OP_PRESSURE_ABS: TRY(op_prep(c, L, true)); break;
    case FLITE_OP_TIMESTEPS: TRY(op_prep(c, L, false)); TRY(op_timesteps(c, L)); break;
    case FLITE_OP_DISSIPATION: TRY(op_dissipation(c, L)); TRY(op_jst(c, L)); break;
    case FLITE_OP_VISCOUS:
      if (!(c->hp.reynolds_number > 0.0)) return fail(FLITE_E_ARG, "viscous term needs ReynoldsNumber > 0");
      TRY(op_prep(c, L, false));
      TRY(op_viscous(c, L));
      break;
    case FLITE_OP_RHS: {
      TRY(op_prep(c, L, false));
      TRY(op_grad(c, L));
      TRY(op_flux(c, L));
      StageArgs S{};
      S.calc_turb = 1;
      TRY(op_stage(c, L, MODE_RHS, S));
      break;
    }
    case FLITE_OP_SA_TIMESTEPS:
      if (!c->P.sa) return fail(FLITE_E_ARG, "needs turbulenceModel=1");
      TRY(op_prep(c, L, false));
      TRY(op_sa_timesteps(c, L));
      break;
    case FLITE_OP_SA_DIFFUSION:
      if (!c->P.sa) return fail(FLITE_E_ARG, "needs turbulenceModel=1");
      TRY(op_prep(c, L, false));
      TRY(op_sa_diffusion(c, L));
      break;
    default: return fail(FLITE_E_ARG, "bad op");
  }
  return sync(c);
}

}  // extern "C"

namespace {
// the monitor's reductions on the device into part (MON_BLOCKS + 1 records
// of MON_REC doubles: the blocks' records, then the total), the total
// copied into tot; res, maxres, maxidx and forces formed from it
int monitor_into(flite_ctx *c, double *part, double res[7], double maxres[7], int maxidx[7], double forces[9]) {
  Level &L = *c->lev[0];
  int nres = c->P.sa ? 6 : 5;
  DevParams P = c->P;
  const int bsel = c->sel;  // selected condition
  const flite_params &hp = c->hp_b[bsel];
  double PI = 4.0 * std::atan(1.0);
  double ar = hp.alpha * PI / 180., br = hp.beta * PI / 180.;
  ForceFrame F;
  F.ft[0] = std::cos(ar) * std::cos(br);
  F.ft[1] = std::sin(br);
  F.ft[2] = std::sin(ar);
  F.fn1[0] = -F.ft[2];
  F.fn1[1] = 0.0;
  F.fn1[2] = F.ft[0];
  F.fn2[0] = -F.ft[1];
  F.fn2[1] = F.ft[0];
  F.fn2[2] = 0.0;
  F.groundplane = hp.groundplane;
  for (int k = 0; k < 3; ++k) F.datum[k] = hp.datum[k];
  int nw = (int)L.wall_dev.size();
  double *total = part + (size_t)MON_REC * MON_BLOCKS;
  TRY(with_level(L, [&](auto V) {
    using R = std::remove_pointer_t<decltype(V.u)>;
    hipLaunchKernelGGL(k_monitor_reduce<R>, dim3(MON_BLOCKS), dim3(NTHR), 0, c->stream, V, P, F, nres, L.host_index,
                       L.wall_nodes, nw, (const R *)L.warea, part, bsel);
    hipLaunchKernelGGL(k_monitor_final, dim3(1), dim3(NTHR), 0, c->stream, (const double *)part, MON_BLOCKS, total);
    return check_launch();
  }));
  double tot[MON_REC];
  HIPTRY(hipMemcpyAsync(tot, total, sizeof(tot), hipMemcpyDeviceToHost, c->stream));
  TRY(sync(c));
  for (int j = 0; j < 7; ++j) { res[j] = 0; maxres[j] = 0; maxidx[j] = 0; }
  for (int j = 0; j < nres; ++j) {
    res[j] = std::sqrt(tot[j]);
    maxres[j] = tot[7 + j];
    maxidx[j] = (tot[7 + j] > 0 ? (int)tot[14 + j] : 0) + c->staged[0].base;  // reference starts at node 1
  }
  const double *f = tot + 21;
  double fct = 2.0 / hp.reference_area;
  forces[0] = fct * f[0];
  forces[1] = fct * f[1];
  forces[2] = fct * f[2];
  forces[3] = fct * f[3];
  forces[4] = f[4];
  forces[5] = f[5];
  forces[6] = fct * f[6];
  forces[7] = fct * f[7];
  forces[8] = 0.0;
  for (int j = 0; j < nres; ++j)
    if (std::isnan(res[j])) return fail(FLITE_E_NAN, "NaN in residual");
  return 0;
}
}  // namespace

extern "C" {

int flite_monitor(flite_ctx *c, double res[7], double maxres[7], int maxidx[7], double forces[9]) {
  TRY(check_ctx(c));
  return monitor_into(c, c->lev[0]->m_part, res, maxres, maxidx, forces);
}

int flite_monitor_noalloc(flite_ctx *c, double res[7], double maxres[7], int maxidx[7], double forces[9]) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  if (!res || !maxres || !maxidx || !forces) return fail(FLITE_E_ARG, "null output array");
  // building the levels or running a due preamble would allocate or launch
  // more than the monitor: the context must be ready
  if (!c->built || c->preamble_due)
    return fail(FLITE_E_STATE, "flite_monitor_noalloc needs a built context (any stepping, field or monitor call "
                               "builds it)");
  if (c->device < 0 || c->device >= SCRATCH_DEVICES || !g_scratch_dev[c->device])
    return fail(FLITE_E_STATE, "no monitor scratch on this context's device");
  // the scratch holds MON_BLOCKS + 1 records whatever the level's size: the
  // reduction's blocks stride over the nodes, so no size exceeds it
  return monitor_into(c, g_scratch_dev[c->device], res, maxres, maxidx, forces);
}


// ---- tangent / adjoint of the level-1 residual ---------------------------
namespace {
int check_linearisable(flite_ctx *c, bool second_order_ok = true, bool laminar_ok = false, bool sa_ok = false) {
  TRY(check_ctx(c));
  const Level &L = *c->lev[0];
  const flite_params &hp = c->hp;
  if (L.fp32) return fail(FLITE_E_UNSUPPORTED, "level 1 must be FP64");
  if (c->nbatch != 1) return fail(FLITE_E_UNSUPPORTED, "tangent/adjoint: single-condition contexts only");
  if (hp.flux_type != 4) return fail(FLITE_E_UNSUPPORTED, "tangent/adjoint: HLLC (fluxType=4) only");
  if (!second_order_ok && L.has_coords && hp.hllc_flux_order > 1)
    return fail(FLITE_E_UNSUPPORTED, "metric derivatives: first-order HLLC (HLLCFluxOrder=1) only");
  if (hp.turbulence_model != 0 && !(sa_ok && hp.turbulence_model == 1))
    return fail(FLITE_E_UNSUPPORTED, sa_ok ? "tangent/adjoint: no turbulence model or Spalart-Allmaras only"
                                           : "this linearisation: no turbulence model (Spalart-Allmaras is "
                                             "linearised in the tangent and adjoint only)");
  if (hp.flow_type == 2) return fail(FLITE_E_UNSUPPORTED, "linearisation: fixed meshes only (no GCL terms)");
  if (!laminar_ok && hp.reynolds_number > 0.0)
    return fail(FLITE_E_UNSUPPORTED, "metric and drag derivatives: inviscid flow (ReynoldsNumber=0) only");
  return 0;
}

bool laminar(const flite_ctx *c) { return c->hp.reynolds_number > 0.0; }

// laminar viscous linearisation scratch: VLIN doubles per node
int alloc_vlin(Level &L) {
  if (!L.vlin) TRY(L.pool.alloc(&L.vlin, (size_t)VLIN * std::max(1, L.n)));
  return 0;
}

// tangent of the viscous term along the direction in V.src, added to V.rhs
// With Spalart-Allmaras it also adds the tangent of the SA diffusion and
// source to component 6 (the eddy viscosity is in the stress already).
// Scratch per node: dq 8 | tau duals 18 | aux duals 6 | grad nu~ duals 8 |
// out 6
int visc_tangent(flite_ctx *c, Level &L) {
  TRY(alloc_vlin(L));
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  const size_t n = L.n;
  double *dq = L.vlin, *dvis = L.vlin + 40 * n;
  ad::Dual *taud = reinterpret_cast<ad::Dual *>(L.vlin + 8 * n);
  ad::Dual *aux = reinterpret_cast<ad::Dual *>(L.vlin + 26 * n);
  ad::Dual *dnud = reinterpret_cast<ad::Dual *>(L.vlin + 32 * n);
  dim3 g(nblk(L.n)), b(NTHR);
  hipLaunchKernelGGL(k_record_tangent, g, b, 0, c->stream, V, P, V.src, dq);
  hipLaunchKernelGGL(k_visc_tangent_node, g, b, 0, c->stream, V, P, (const double *)dq, taud, aux);
  hipLaunchKernelGGL(k_visc_tangent_flux, g, b, 0, c->stream, V, P, (const double *)dq, (const ad::Dual *)taud, dvis);
  if (P.sa) {
    hipLaunchKernelGGL(k_sa_diff1_tangent, g, b, 0, c->stream, V, P, (const double *)dq, dnud);
    hipLaunchKernelGGL(k_sa_tangent, g, b, 0, c->stream, V, P, (const double *)dq, (const ad::Dual *)aux,
                       (const ad::Dual *)dnud, dvis + 5 * n);
  }
  hipLaunchKernelGGL(k_add_vector, g, b, 0, c->stream, V.rhs, (const double *)dvis, L.n, P.sa ? 6 : 5);
  return check_launch();
}

// the viscous linearisation's scratch (VLIN doubles per node): stress duals
// 18 | taub 9 | ubv 3 | gbar 12 | tbar 1 | abar 4 | (rho, nu~) shares 2 |
// grad nu~ shares 3 | SA (d, v) 4 | mesh shares 5 | friction area 1 and
// normal 3 shares
struct Vlin {
  ad::Dual *taud;
  double *taub, *ubv, *gbar, *tbar, *abar, *rnbar, *gnbar, *sad, *mbar, *gaf, *gnf;
  explicit Vlin(Level &L) {
    const size_t n = L.n;
    double *b = L.vlin;
    taud = reinterpret_cast<ad::Dual *>(b);
    taub = b + 18 * n;
    ubv = b + 27 * n;
    gbar = b + 30 * n;
    tbar = b + 42 * n;
    abar = b + 43 * n;
    rnbar = b + 47 * n;
    gnbar = b + 49 * n;
    sad = b + 52 * n;
    mbar = b + 56 * n;
    gaf = b + 61 * n;
    gnf = b + 62 * n;
  }
};

// adjoint of the viscous term against the multipliers in V.src, added to V.rhs
// With Spalart-Allmaras it adds the transposes of the eddy viscosity in the
// stress, the SA source and the SA diffusion (multipliers lam(6)).
int visc_adjoint(flite_ctx *c, Level &L) {
  TRY(alloc_vlin(L));
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  Vlin S(L);
  dim3 g(nblk(L.n)), b(NTHR);
  hipLaunchKernelGGL(k_visc_tangent_node, g, b, 0, c->stream, V, P, (const double *)nullptr, S.taud, (ad::Dual *)nullptr);
  hipLaunchKernelGGL(k_visc_adjoint_tau, g, b, 0, c->stream, V, P, V.src, (const ad::Dual *)S.taud, S.taub, S.ubv);
  hipLaunchKernelGGL(k_visc_adjoint_node, g, b, 0, c->stream, V, P, V.src, (const double *)S.taub, S.gbar, S.tbar,
                     S.rnbar, (double *)nullptr);
  if (P.sa)
    hipLaunchKernelGGL(k_sa_diff_adjoint1, g, b, 0, c->stream, V, P, V.src, S.gnbar, S.tbar, S.rnbar, (double *)nullptr,
                       (double *)nullptr);
  hipLaunchKernelGGL(k_visc_adjoint_grad, g, b, 0, c->stream, V, (const double *)S.gbar, (const double *)S.tbar,
                     (const double *)S.ubv, S.abar);
  hipLaunchKernelGGL(k_visc_adjoint_record, g, b, 0, c->stream, V, P, (const double *)S.abar, V.rhs);
  if (P.sa)
    hipLaunchKernelGGL(k_sa_diff_adjoint2, g, b, 0, c->stream, V, P, (const double *)S.gnbar, (const double *)S.rnbar,
                       V.rhs);
  return check_launch();
}

// per-side and per-boundary-side mesh shares of the viscous terms (device
// side order), allocated on first use
int alloc_vmesh(Level &L) {
  if (!L.vmesh) TRY(L.pool.alloc(&L.vmesh, 3 * ((size_t)std::max(1, L.ne) + (size_t)std::max(1, L.nbs))));
  return 0;
}

// d(lam^T R_visc)/d(mesh) for lam in V.src: side weights into L.vmesh
// (3 per side), boundary-side weights after them (3 per boundary side), and
// per node in S.mbar: volume, wall normal, wall distance.  Records current.
int visc_mesh_adjoint(flite_ctx *c, Level &L) {
  TRY(alloc_vlin(L));
  TRY(alloc_vmesh(L));
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  Vlin S(L);
  double *dgw = L.vmesh, *dgwb = L.vmesh + 3 * (size_t)std::max(1, L.ne);
  dim3 g(nblk(L.n)), b(NTHR);
  const size_t ns = 3 * ((size_t)std::max(1, L.ne) + (size_t)std::max(1, L.nbs));
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk((long long)ns)), dim3(NTHR), 0, c->stream, L.vmesh, 0.0, ns);
  hipLaunchKernelGGL(k_visc_tangent_node, g, b, 0, c->stream, V, P, (const double *)nullptr, S.taud, (ad::Dual *)nullptr);
  hipLaunchKernelGGL(k_visc_adjoint_tau, g, b, 0, c->stream, V, P, V.src, (const ad::Dual *)S.taud, S.taub, S.ubv);
  hipLaunchKernelGGL(k_visc_adjoint_node, g, b, 0, c->stream, V, P, V.src, (const double *)S.taub, S.gbar, S.tbar,
                     S.rnbar, S.mbar);
  if (P.sa)
    hipLaunchKernelGGL(k_sa_diff_adjoint1, g, b, 0, c->stream, V, P, V.src, S.gnbar, S.tbar, S.rnbar, S.mbar, S.sad);
  if (L.ne > 0)
    hipLaunchKernelGGL(k_visc_mesh_side, dim3(nblk(L.ne)), b, 0, c->stream, V, P, V.src, (const ad::Dual *)S.taud,
                       (const double *)S.gbar, (const double *)S.gnbar, (const double *)S.sad, dgw);
  if (L.nbs > 0)
    hipLaunchKernelGGL(k_visc_mesh_bside, dim3(nblk(L.nbs)), b, 0, c->stream, V, P, V.src, (const ad::Dual *)S.taud,
                       (const double *)S.gbar, (const double *)S.gnbar, (const double *)S.sad, dgwb);
  return check_launch();
}

// the friction term of the drag (ReynoldsNumber > 0) reversed: its shares
// on the stress records through the node pass (lam null) into S.gbar,
// S.tbar, S.rnbar and, with mesh, S.mbar; the direct area and normal shares
// in S.gaf, S.gnf.  Records current.
int visc_friction_reverse(flite_ctx *c, Level &L, bool mesh) {
  TRY(alloc_vlin(L));
  DevParams P = c->P;
  const flite_params &hp = c->hp;
  LevelDev<double> V = L.view<double>();
  Vlin S(L);
  const size_t n = L.n;
  dim3 g(nblk(L.n)), b(NTHR);
  double PI = 4.0 * std::atan(1.0);
  double ar = hp.alpha * PI / 180., br = hp.beta * PI / 180.;
  double t0 = std::cos(ar) * std::cos(br), t1 = std::sin(br), t2 = std::sin(ar);
  hipLaunchKernelGGL(k_visc_tangent_node, g, b, 0, c->stream, V, P, (const double *)nullptr, S.taud, (ad::Dual *)nullptr);
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk((long long)(9 * n))), b, 0, c->stream, S.taub, 0.0, 9 * n);
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk((long long)(4 * n))), b, 0, c->stream, S.gaf, 0.0, 4 * n);
  const int nw = (int)L.wall_dev.size();
  if (nw > 0)
    hipLaunchKernelGGL(k_friction_shares, dim3(nblk(nw)), b, 0, c->stream, V, (const int *)L.wall_nodes, nw,
                       (const double *)L.warea, 2.0 / hp.reference_area, t0, t1, t2, hp.groundplane,
                       (const ad::Dual *)S.taud, S.taub, S.gaf, S.gnf);
  hipLaunchKernelGGL(k_visc_adjoint_node, g, b, 0, c->stream, V, P, (const double *)nullptr, (const double *)S.taub,
                     S.gbar, S.tbar, S.rnbar, mesh ? S.mbar : (double *)nullptr);
  return check_launch();
}

// the wall stress of the current state (the viscous node pass), for the
// drag's friction term
int wall_stress(flite_ctx *c, Level &L) {
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  launch_gather<ViscNodePart<double, true>, NoPart<double>, NoPart<double>>(c, L, V, P);
  return check_launch();
}

// host (N, ncomp) column-major -> level-1 device array src (6 components;
// ncomp = 5, or 6 with nu~ as the sixth)
int put_vector(flite_ctx *c, const double *x, int ld, int ncomp = 5) {
  Level &L = *c->lev[0];
  if (!x || ld < L.n) return fail(FLITE_E_ARG, "bad input vector");
  std::vector<double> h(6 * (size_t)L.n, 0.0);
  for (int k = 0; k < ncomp; ++k)
    for (int i = 0; i < L.n; ++i) h[(size_t)k * L.n + i] = x[(size_t)k * ld + i];
  return flite_field_set(c, 1, FLITE_FIELD_SOURCE, h.data(), L.n);
}
int put_vector5(flite_ctx *c, const double *x, int ld) { return put_vector(c, x, ld, 5); }

int get_vector(flite_ctx *c, double *y, int ld, int ncomp = 5) {
  Level &L = *c->lev[0];
  if (!y || ld < L.n) return fail(FLITE_E_ARG, "bad output vector");
  std::vector<double> h(6 * (size_t)L.n);
  TRY(flite_field_get(c, 1, FLITE_FIELD_RHS, h.data(), L.n));
  for (int k = 0; k < ncomp; ++k)
    for (int i = 0; i < L.n; ++i) y[(size_t)k * ld + i] = h[(size_t)k * L.n + i];
  return 0;
}
int get_vector5(flite_ctx *c, double *y, int ld) { return get_vector(c, y, ld, 5); }

bool second_order(const flite_ctx *c, const Level &L) {
  return L.has_coords && c->hp.hllc_flux_order > 1 && c->hp.flux_type == 4;
}

// second-order linearisation scratch: per node DQ + 52 doubles (tangent: dq
// DQ + dual (G, B) 52; adjoint: state shares 5, (G, B) shares 25, record
// shares 5)
// and 10 extremal-neighbour indices
int alloc_lin(Level &L) {
  if (!L.lin) TRY(L.pool.alloc(&L.lin, (DQ + 52) * (size_t)std::max(1, L.n)));
  if (!L.amin) TRY(L.pool.alloc(&L.amin, 10 * (size_t)std::max(1, L.n)));
  return 0;
}
}  // namespace

namespace {
// (dR/dU) V.src into V.rhs on level 1 (src filled by the caller)
int tangent_core(flite_ctx *c) {
  Level &L = *c->lev[0];
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  if (second_order(c, L)) {
    // records, gradients and bounds of the current state, then dq,
    // (dG, dB) and the side fluxes' tangents
    TRY(op_prep(c, L, true));
    TRY(op_grad(c, L));
    TRY(alloc_lin(L));
    double *dq = L.lin;
    ad::Dual *gd = reinterpret_cast<ad::Dual *>(L.lin + DQ * (size_t)L.n);
    hipLaunchKernelGGL(k_record_tangent, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, P, V.src, dq);
    hipLaunchKernelGGL(k_grad_tangent, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, (const double *)dq, gd);
    if (L.ne > 0) {
      dim3 ge(nblk(L.ne)), be(NTHR);
      const double *dqc = dq;
      const ad::Dual *gdc = gd;
      if (P.nsig == 4 && P.sa) hipLaunchKernelGGL((k_flux_tangent2<4, true>), ge, be, 0, c->stream, V, P, dqc, gdc);
      else if (P.nsig == 4) hipLaunchKernelGGL((k_flux_tangent2<4, false>), ge, be, 0, c->stream, V, P, dqc, gdc);
      else if (P.sa) hipLaunchKernelGGL((k_flux_tangent2<2, true>), ge, be, 0, c->stream, V, P, dqc, gdc);
      else hipLaunchKernelGGL((k_flux_tangent2<2, false>), ge, be, 0, c->stream, V, P, dqc, gdc);
    }
  } else if (L.ne > 0) {
    if (P.sa) hipLaunchKernelGGL(k_flux_tangent<true>, dim3(nblk(L.ne)), dim3(NTHR), 0, c->stream, V, P, V.src);
    else hipLaunchKernelGGL(k_flux_tangent<false>, dim3(nblk(L.ne)), dim3(NTHR), 0, c->stream, V, P, V.src);
  }
  L.flux_partials = false;  // the tangent fluxes are per-side records
  if (L.nbs > 0) hipLaunchKernelGGL(k_bflux_tangent, dim3(nblk(L.nbs)), dim3(NTHR), 0, c->stream, V, P, V.src);
  TRY(check_launch());
  StageArgs S{};
  TRY(op_stage(c, L, MODE_RHS, S));
  if (laminar(c)) {
    if (!second_order(c, L)) TRY(op_prep(c, L, true));
    TRY(visc_tangent(c, L));
  }
  return 0;
}

// (dR/dU)^T V.src into V.rhs on level 1 (src filled by the caller)
int adjoint_core(flite_ctx *c) {
  Level &L = *c->lev[0];
  if (second_order(c, L)) {
    // records, gradients and bounds of the current state, the bounds'
    // extremal neighbours, then the per-side shares and their transposes
    TRY(op_prep(c, L, true));
    TRY(op_grad(c, L));
    TRY(alloc_lin(L));
    if (!L.abuf2) TRY(L.pool.alloc(&L.abuf2, (size_t)LIN_SIDE * std::max(1, L.ne)));
    DevParams P = c->P;
    LevelDev<double> V = L.view<double>();
    double *ub = L.lin, *gbn = L.lin + 5 * (size_t)L.n, *qb = L.lin + 30 * (size_t)L.n;
    hipLaunchKernelGGL(k_minmax_arg, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, L.amin);
    if (L.ne > 0) {
      if (P.nsig == 4)
        hipLaunchKernelGGL(k_flux_adjoint2<4>, dim3(nblk(L.ne)), dim3(NTHR), 0, c->stream, V, P, V.src, L.abuf2);
      else
        hipLaunchKernelGGL(k_flux_adjoint2<2>, dim3(nblk(L.ne)), dim3(NTHR), 0, c->stream, V, P, V.src, L.abuf2);
    }
    if (L.nbs > 0) hipLaunchKernelGGL(k_bflux_adjoint, dim3(nblk(L.nbs)), dim3(NTHR), 0, c->stream, V, P, V.src);
    hipLaunchKernelGGL(k_adjoint_gather2, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, (const double *)L.abuf2, ub, gbn);
    hipLaunchKernelGGL(k_grad_adjoint, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, (const double *)gbn,
                       (const int *)L.amin, qb);
    hipLaunchKernelGGL(k_record_adjoint, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, P, (const double *)ub,
                       (const double *)qb, V.rhs);
    if (P.sa) hipLaunchKernelGGL(k_sa_conv_adjoint, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, P, V.src, V.rhs);
    TRY(check_launch());
    if (laminar(c)) TRY(visc_adjoint(c, L));
    return 0;
  }
  if (!L.abuf) TRY(L.pool.alloc(&L.abuf, 10 * (size_t)std::max(1, L.ne)));
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  if (L.ne > 0) hipLaunchKernelGGL(k_flux_adjoint, dim3(nblk(L.ne)), dim3(NTHR), 0, c->stream, V, P, V.src, L.abuf);
  if (L.nbs > 0) hipLaunchKernelGGL(k_bflux_adjoint, dim3(nblk(L.nbs)), dim3(NTHR), 0, c->stream, V, P, V.src);
  hipLaunchKernelGGL(k_adjoint_gather, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, L.abuf, V.rhs);
  if (P.sa) hipLaunchKernelGGL(k_sa_conv_adjoint, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, V, P, V.src, V.rhs);
  TRY(check_launch());
  if (laminar(c)) {
    TRY(op_prep(c, L, true));
    TRY(visc_adjoint(c, L));
  }
  return 0;
}

// device-vector form of an apply: in and out on the context's GPU
int apply_device(flite_ctx *c, const double *in, double *out, int (*core)(flite_ctx *)) {
  TRY(check_linearisable(c, true, true, true));
  if (!in || !out) return fail(FLITE_E_ARG, "null vector");
  for (const void *p : {(const void *)in, (const void *)out}) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeDevice || a.device != c->device) {
      (void)hipGetLastError();
      return fail(FLITE_E_ARG, "vectors must be device memory on the context's GPU");
    }
  }
  Level &L = *c->lev[0];
  const int nc = c->P.sa ? 6 : 5;
  LevelDev<double> V = L.view<double>();
  // the producer of `in` may run on another stream: wait for the device
  HIPTRY(hipDeviceSynchronize());
  hipLaunchKernelGGL(k_load_host_order, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, in, (const int *)L.host_index,
                     L.n, nc, V.src);
  TRY(check_launch());
  TRY(core(c));
  hipLaunchKernelGGL(k_store_host_order, dim3(nblk(L.n)), dim3(NTHR), 0, c->stream, (const double *)V.rhs,
                     (const int *)L.host_index, L.n, nc, out);
  TRY(check_launch());
  HIPTRY(hipStreamSynchronize(c->stream));
  return 0;
}
}  // namespace

int flite_tangent_apply(flite_ctx *c, const double *du, int ld_du, double *out, int ld_out) {
  TRY(check_linearisable(c, true, true, true));
  const int nc = c->P.sa ? 6 : 5;
  TRY(put_vector(c, du, ld_du, nc));
  TRY(tangent_core(c));
  return get_vector(c, out, ld_out, nc);
}

int flite_adjoint_apply(flite_ctx *c, const double *lam, int ld_lam, double *out, int ld_out) {
  TRY(check_linearisable(c, true, true, true));
  const int nc = c->P.sa ? 6 : 5;
  TRY(put_vector(c, lam, ld_lam, nc));
  TRY(adjoint_core(c));
  return get_vector(c, out, ld_out, nc);
}

int flite_tangent_apply_device(flite_ctx *c, const double *du, double *out) {
  return apply_device(c, du, out, tangent_core);
}

int flite_adjoint_apply_device(flite_ctx *c, const double *lam, double *out) {
  return apply_device(c, lam, out, adjoint_core);
}

int flite_drag_state_gradient(flite_ctx *c, double *drag, double *dJdU, int ld) {
  TRY(check_linearisable(c, true, true, true));
  Level &L = *c->lev[0];
  if (!dJdU || ld < L.n) return fail(FLITE_E_ARG, "bad output vector");
  const flite_params &hp = c->hp;
  TRY(op_prep(c, L, true));
  if (laminar(c)) TRY(wall_stress(c, L));
  double res[7], mx[7], forces[9];
  int idx[7];
  TRY(flite_monitor(c, res, mx, idx, forces));
  if (drag) *drag = forces[1];
  double PI = 4.0 * std::atan(1.0);
  double ar = hp.alpha * PI / 180., br = hp.beta * PI / 180.;
  double t0 = std::cos(ar) * std::cos(br), t1 = std::sin(br), t2 = std::sin(ar);
  LevelDev<double> V = L.view<double>();
  size_t cnt = 6 * (size_t)L.n;
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk((long long)cnt)), dim3(NTHR), 0, c->stream, V.rhs, 0.0, cnt);
  int nw = (int)L.wall_dev.size();
  if (nw > 0)
    hipLaunchKernelGGL(k_drag_state_gradient, dim3(nblk(nw)), dim3(NTHR), 0, c->stream, V, c->P, L.wall_nodes, nw,
                       (const double *)L.warea, 2.0 / hp.reference_area, t0, t1, t2, hp.groundplane, V.rhs);
  TRY(check_launch());
  if (laminar(c)) {
    // the friction term through the wall stress: its shares on the stress
    // records, then the Green-Gauss transpose and the records' dependence
    // on the state (and the eddy viscosity's on rho and nu~)
    TRY(visc_friction_reverse(c, L, false));
    Vlin S(L);
    dim3 g(nblk(L.n)), b(NTHR);
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk((long long)(3 * (size_t)L.n))), b, 0, c->stream, S.ubv, 0.0,
                       3 * (size_t)L.n);
    hipLaunchKernelGGL(k_visc_adjoint_grad, g, b, 0, c->stream, V, (const double *)S.gbar, (const double *)S.tbar,
                       (const double *)S.ubv, S.abar);
    hipLaunchKernelGGL(k_visc_adjoint_record, g, b, 0, c->stream, V, c->P, (const double *)S.abar, V.rhs);
    if (c->P.sa) hipLaunchKernelGGL(k_add_rho_nu, g, b, 0, c->stream, (const double *)S.rnbar, L.n, V.rhs);
    TRY(check_launch());
  }
  return get_vector(c, dJdU, ld, c->P.sa ? 6 : 5);
}

namespace {
// the chain from per-node shares on the wall areas A_i = sum 4 |w_b| and the
// unit normals n_i = m_i / |m_i|, m_i = -sum w_b (sums over the boundary
// sides at the node, as build_level forms warea and nrm; host numbering) to
// the boundary-side weights, added to gwb
void wb_chain(const Staged &S, int n, int nbs, const std::vector<double> &gA, const std::vector<double> &gN,
              double *gwb, int ld_wb) {
  std::vector<double> m(3 * (size_t)n, 0.0), gM(3 * (size_t)n, 0.0);
  for (int s = 0; s < nbs; ++s)
    for (int q = 0; q < 3; ++q) {
      const double w = S.wb[(size_t)q * nbs + s];
      m[3 * (size_t)S.b1[s] + q] -= w;
      m[3 * (size_t)S.b2[s] + q] -= w;
    }
  for (int i = 0; i < n; ++i) {
    const double *gn = &gN[3 * (size_t)i], *mi = &m[3 * (size_t)i];
    if (gn[0] == 0.0 && gn[1] == 0.0 && gn[2] == 0.0) continue;
    const double mn = std::sqrt(mi[0] * mi[0] + mi[1] * mi[1] + mi[2] * mi[2]);
    if (!(mn > 1e-10)) continue;  // zero normal: no dependence
    const double nh[3] = {mi[0] / mn, mi[1] / mn, mi[2] / mn};
    const double gNn = gn[0] * nh[0] + gn[1] * nh[1] + gn[2] * nh[2];
    for (int q = 0; q < 3; ++q) gM[3 * (size_t)i + q] = (gn[q] - gNn * nh[q]) / mn;
  }
  for (int s = 0; s < nbs; ++s) {
    double w[3], a = 0;
    for (int q = 0; q < 3; ++q) { w[q] = S.wb[(size_t)q * nbs + s]; a += w[q] * w[q]; }
    a = std::sqrt(a);
    double g[3] = {0, 0, 0};
    for (int i : {S.b1[s], S.b2[s]}) {
      for (int q = 0; q < 3; ++q) g[q] -= gM[3 * (size_t)i + q];
      if (a > 0)
        for (int q = 0; q < 3; ++q) g[q] += gA[i] * 4.0 * w[q] / a;
    }
    for (int q = 0; q < 3; ++q) gwb[(size_t)q * ld_wb + s] += g[q];
  }
}

// the first-order convective residual's side and boundary-side weight
// shares for the multipliers in src, into gw and gwb (host layouts)
int metric_adjoint_conv(flite_ctx *c, double *gw, int ld_w, double *gwb, int ld_wb) {
  Level &L = *c->lev[0];
  if (!L.abuf) TRY(L.pool.alloc(&L.abuf, 10 * (size_t)std::max(1, L.ne)));
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  // side shares into abuf (3 per side), boundary shares into bbuf (3 per side)
  if (L.ne > 0) hipLaunchKernelGGL(k_flux_metric_adjoint, dim3(nblk(L.ne)), dim3(NTHR), 0, c->stream, V, P, V.src, L.abuf);
  if (L.nbs > 0)
    hipLaunchKernelGGL(k_bflux_metric_adjoint, dim3(nblk(L.nbs)), dim3(NTHR), 0, c->stream, V, P, V.src, V.bbuf);
  TRY(check_launch());
  std::vector<double> hw(3 * (size_t)L.ne), hb(3 * (size_t)L.nbs);
  if (L.ne > 0) HIPTRY(hipMemcpyAsync(hw.data(), L.abuf, hw.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (L.nbs > 0) HIPTRY(hipMemcpyAsync(hb.data(), V.bbuf, hb.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < L.ne; ++k)
    for (int q = 0; q < 3; ++q) gw[(size_t)q * ld_w + L.side_host[k]] = hw[3 * (size_t)k + q];
  for (int s = 0; s < L.nbs; ++s)
    for (int q = 0; q < 3; ++q) gwb[(size_t)q * ld_wb + s] = hb[3 * (size_t)s + q];
  return 0;
}

// the viscous and SA terms' mesh shares for the multipliers in src, added
// to gw, gwb and gvol; gwd (may be null without SA) receives the wall
// distance shares.  The wall normals' shares go to gwb through wb_chain.
int mesh_adjoint_visc(flite_ctx *c, double *gw, int ld_w, double *gwb, int ld_wb, double *gvol, double *gwd) {
  Level &L = *c->lev[0];
  const int n = L.n, ne = L.ne, nbs = L.nbs;
  TRY(op_prep(c, L, true));
  TRY(visc_mesh_adjoint(c, L));
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  double *dgw = L.vmesh, *dgwb = L.vmesh + 3 * (size_t)std::max(1, ne);
  if (P.sa && ne > 0)
    hipLaunchKernelGGL(k_sa_conv_mesh_adjoint, dim3(nblk(ne)), dim3(NTHR), 0, c->stream, V, P, V.src, dgw);
  TRY(check_launch());
  Vlin S(L);
  std::vector<double> hw(3 * (size_t)ne), hb(3 * (size_t)nbs), hm(5 * (size_t)n);
  if (ne > 0) HIPTRY(hipMemcpyAsync(hw.data(), dgw, hw.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (nbs > 0) HIPTRY(hipMemcpyAsync(hb.data(), dgwb, hb.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipMemcpyAsync(hm.data(), S.mbar, hm.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < ne; ++k)
    for (int q = 0; q < 3; ++q) gw[(size_t)q * ld_w + L.side_host[k]] += hw[3 * (size_t)k + q];
  for (int s = 0; s < nbs; ++s)
    for (int q = 0; q < 3; ++q) gwb[(size_t)q * ld_wb + s] += hb[3 * (size_t)s + q];
  std::vector<double> gA(n, 0.0), gN(3 * (size_t)n, 0.0);
  for (int hh = 0; hh < n; ++hh) {
    const int d = L.h2d[hh];
    gvol[hh] += hm[5 * (size_t)d];
    if (gwd) gwd[hh] = hm[5 * (size_t)d + 4];
    for (int q = 0; q < 3; ++q) gN[3 * (size_t)hh + q] = hm[5 * (size_t)d + 1 + q];
  }
  wb_chain(c->staged[0], n, nbs, gA, gN, gwb, ld_wb);
  return 0;
}
}  // namespace

int flite_metric_adjoint(flite_ctx *c, const double *lam, int ld_lam, double *gw, int ld_w, double *gwb,
                         int ld_wb) {
  TRY(check_linearisable(c, false));
  Level &L = *c->lev[0];
  if (!gw || ld_w < L.ne || !gwb || ld_wb < L.nbs) return fail(FLITE_E_ARG, "bad output arrays");
  TRY(put_vector5(c, lam, ld_lam));
  return metric_adjoint_conv(c, gw, ld_w, gwb, ld_wb);
}

int flite_mesh_adjoint(flite_ctx *c, const double *lam, int ld_lam, double *gw, int ld_w, double *gwb, int ld_wb,
                       double *gvol, double *gxyz, int ld_x, double *gwd) {
  TRY(check_linearisable(c, true, true, true));
  Level &L = *c->lev[0];
  if (!gw || ld_w < L.ne || !gwb || ld_wb < L.nbs || !gvol || !gxyz || ld_x < L.n)
    return fail(FLITE_E_ARG, "bad output arrays");
  if (c->P.sa && !gwd) return fail(FLITE_E_ARG, "Spalart-Allmaras: gwd (wall distance shares) is required");
  const int n = L.n, ne = L.ne, nbs = L.nbs;
  const int nc = c->P.sa ? 6 : 5;
  if (gwd)
    for (int i = 0; i < n; ++i) gwd[i] = 0.0;
  if (!second_order(c, L)) {
    // the first-order convective residual depends on the side weights only
    TRY(put_vector(c, lam, ld_lam, nc));
    TRY(metric_adjoint_conv(c, gw, ld_w, gwb, ld_wb));
    for (int i = 0; i < n; ++i) gvol[i] = 0.0;
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < n; ++i) gxyz[(size_t)k * ld_x + i] = 0.0;
    if (laminar(c)) TRY(mesh_adjoint_visc(c, gw, ld_w, gwb, ld_wb, gvol, gwd));
    return 0;
  }
  // records, gradients and bounds of the state; the state adjoint's side
  // shares give d(lam^T R)/d(G, B) per node (gbn); then the side, boundary
  // side and volume shares
  TRY(op_prep(c, L, true));
  TRY(op_grad(c, L));
  TRY(alloc_lin(L));
  if (!L.abuf2) TRY(L.pool.alloc(&L.abuf2, (size_t)LIN_SIDE * std::max(1, ne)));
  if (!L.abuf) TRY(L.pool.alloc(&L.abuf, 10 * (size_t)std::max(1, ne)));
  TRY(put_vector(c, lam, ld_lam, nc));
  DevParams P = c->P;
  LevelDev<double> V = L.view<double>();
  double *ub = L.lin, *gbn = L.lin + 5 * (size_t)n, *gv = L.lin + 30 * (size_t)n;
  double *dgw = L.abuf, *dgdx = L.abuf + 3 * (size_t)ne;
  if (ne > 0) {
    if (P.nsig == 4)
      hipLaunchKernelGGL(k_flux_adjoint2<4>, dim3(nblk(ne)), dim3(NTHR), 0, c->stream, V, P, V.src, L.abuf2);
    else
      hipLaunchKernelGGL(k_flux_adjoint2<2>, dim3(nblk(ne)), dim3(NTHR), 0, c->stream, V, P, V.src, L.abuf2);
  }
  if (nbs > 0) hipLaunchKernelGGL(k_bflux_adjoint, dim3(nblk(nbs)), dim3(NTHR), 0, c->stream, V, P, V.src);
  hipLaunchKernelGGL(k_adjoint_gather2, dim3(nblk(n)), dim3(NTHR), 0, c->stream, V, (const double *)L.abuf2, ub, gbn);
  if (ne > 0) {
    if (P.nsig == 4)
      hipLaunchKernelGGL(k_flux_mesh_adjoint2<4>, dim3(nblk(ne)), dim3(NTHR), 0, c->stream, V, P, V.src,
                         (const double *)gbn, dgw, dgdx);
    else
      hipLaunchKernelGGL(k_flux_mesh_adjoint2<2>, dim3(nblk(ne)), dim3(NTHR), 0, c->stream, V, P, V.src,
                         (const double *)gbn, dgw, dgdx);
  }
  if (nbs > 0) {
    hipLaunchKernelGGL(k_bflux_metric_adjoint, dim3(nblk(nbs)), dim3(NTHR), 0, c->stream, V, P, V.src, V.bbuf);
    hipLaunchKernelGGL(k_bgrad_mesh_adjoint, dim3(nblk(nbs)), dim3(NTHR), 0, c->stream, V, (const double *)gbn,
                       V.bbuf);
  }
  hipLaunchKernelGGL(k_vol_mesh_adjoint, dim3(nblk(n)), dim3(NTHR), 0, c->stream, V, (const double *)gbn, gv);
  TRY(check_launch());
  std::vector<double> hw(3 * (size_t)ne), hdx(3 * (size_t)ne), hb(3 * (size_t)nbs), hv(n);
  std::vector<int> ea(ne), eb(ne);
  if (ne > 0) {
    HIPTRY(hipMemcpyAsync(hw.data(), dgw, hw.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(hdx.data(), dgdx, hdx.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(ea.data(), L.ea, ne * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(eb.data(), L.eb, ne * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  }
  if (nbs > 0) HIPTRY(hipMemcpyAsync(hb.data(), V.bbuf, hb.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipMemcpyAsync(hv.data(), gv, hv.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < ne; ++k)
    for (int q = 0; q < 3; ++q) gw[(size_t)q * ld_w + L.side_host[k]] = hw[3 * (size_t)k + q];
  for (int s = 0; s < nbs; ++s)
    for (int q = 0; q < 3; ++q) gwb[(size_t)q * ld_wb + s] = hb[3 * (size_t)s + q];
  // dx = X_b - X_a: its share goes to +b and -a (device numbering)
  std::vector<double> gx(3 * (size_t)n, 0.0);
  for (int k = 0; k < ne; ++k)
    for (int q = 0; q < 3; ++q) {
      gx[3 * (size_t)eb[k] + q] += hdx[3 * (size_t)k + q];
      gx[3 * (size_t)ea[k] + q] -= hdx[3 * (size_t)k + q];
    }
  for (int hh = 0; hh < n; ++hh) {
    const int d = L.h2d[hh];
    gvol[hh] = hv[d];
    for (int q = 0; q < 3; ++q) gxyz[(size_t)q * ld_x + hh] = gx[3 * (size_t)d + q];
  }
  if (laminar(c)) TRY(mesh_adjoint_visc(c, gw, ld_w, gwb, ld_wb, gvol, gwd));
  return 0;
}

namespace {
// the drag at the current state (records and, with ReynoldsNumber > 0, the
// wall stress recomputed first), and its pressure term's shares on the
// counted wall nodes' areas and unit normals (host numbering)
int drag_pressure_shares(flite_ctx *c, double *drag, std::vector<double> &gA, std::vector<double> &gN) {
  Level &L = *c->lev[0];
  const Staged &S = c->staged[0];
  const flite_params &hp = c->hp;
  TRY(op_prep(c, L, true));
  if (laminar(c)) TRY(wall_stress(c, L));
  double res[7], mx[7], forces[9];
  int idx[7];
  TRY(flite_monitor(c, res, mx, idx, forces));
  if (drag) *drag = forces[1];
  const int n = L.n, nbs = L.nbs;
  std::vector<double> u(6 * (size_t)n);
  TRY(flite_state_get(c, 1, u.data(), n, 5, 0));
  double PI = 4.0 * std::atan(1.0);
  double ar = hp.alpha * PI / 180., br = hp.beta * PI / 180.;
  const double t[3] = {std::cos(ar) * std::cos(br), std::sin(br), std::sin(ar)};
  const double f = 2.0 / hp.reference_area, g1 = hp.gamma - 1.0;
  const double pinf = 1.0 / (hp.gamma * hp.mach_number * hp.mach_number);
  std::vector<double> m(3 * (size_t)n, 0.0), area(n, 0.0);
  for (int s = 0; s < nbs; ++s) {
    double w[3], a = 0;
    for (int q = 0; q < 3; ++q) { w[q] = S.wb[(size_t)q * nbs + s]; a += w[q] * w[q]; }
    a = 4.0 * std::sqrt(a);
    for (int q = 0; q < 3; ++q) { m[3 * (size_t)S.b1[s] + q] -= w[q]; m[3 * (size_t)S.b2[s] + q] -= w[q]; }
    area[S.b1[s]] += a;
    area[S.b2[s]] += a;
  }
  gA.assign(n, 0.0);
  gN.assign(3 * (size_t)n, 0.0);
  for (int d : L.wall_dev) {
    int i = L.d2h[d];
    double z = S.xyz.empty() ? 0.0 : S.xyz[2 * (size_t)n + i];
    if (!(z > hp.groundplane + 0.01)) continue;
    double r = u[i], m1 = u[(size_t)n + i], m2 = u[2 * (size_t)n + i], m3 = u[3 * (size_t)n + i],
           E = u[4 * (size_t)n + i];
    double p = g1 * (E - 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) / r);
    p = p > 0.1 ? p : 0.1;
    const double *mi = &m[3 * (size_t)i];
    double mn = std::sqrt(mi[0] * mi[0] + mi[1] * mi[1] + mi[2] * mi[2]);
    if (!(mn > 1e-10)) continue;  // zero normal: no dependence
    double nt = (mi[0] * t[0] + mi[1] * t[1] + mi[2] * t[2]) / mn;
    gA[i] = -f * (p - pinf) * nt;
    for (int q = 0; q < 3; ++q) gN[3 * (size_t)i + q] = -f * area[i] * (p - pinf) * t[q];
  }
  return 0;
}
}  // namespace

// pressure drag of calculateLiftAndDrag (DEF:14757-14803) as a function of
// the boundary-side weights at a fixed state: the wall node areas
// A_i = sum 4|w_b| and normals n_i = m_i/|m_i|, m_i = -sum w_b, over the
// boundary sides at the node (as build_level forms warea and nrm)
int flite_drag_metric_gradient(flite_ctx *c, double *drag, double *gwb, int ld_wb) {
  TRY(check_linearisable(c));  // the drag itself does not depend on the residual's order
  Level &L = *c->lev[0];
  if (!gwb || ld_wb < L.nbs) return fail(FLITE_E_ARG, "bad output array");
  std::vector<double> gA, gN;
  TRY(drag_pressure_shares(c, drag, gA, gN));
  for (int q = 0; q < 3; ++q)
    for (int s = 0; s < L.nbs; ++s) gwb[(size_t)q * ld_wb + s] = 0.0;
  wb_chain(c->staged[0], L.n, L.nbs, gA, gN, gwb, ld_wb);
  return 0;
}

// the drag (pressure and, with ReynoldsNumber > 0, friction) as a function
// of the side weights, boundary-side weights and volumes at a fixed state.
// The friction fct A_i ws_i.t reads the wall areas and normals and the wall
// stress, whose Green-Gauss sums read the weights and whose gradients read
// the volumes.
int flite_drag_mesh_gradient(flite_ctx *c, double *drag, double *gw, int ld_w, double *gwb, int ld_wb, double *gvol) {
  TRY(check_linearisable(c, true, true, true));
  Level &L = *c->lev[0];
  if (!gw || ld_w < L.ne || !gwb || ld_wb < L.nbs || !gvol) return fail(FLITE_E_ARG, "bad output arrays");
  const int n = L.n, ne = L.ne, nbs = L.nbs;
  std::vector<double> gA, gN;
  TRY(drag_pressure_shares(c, drag, gA, gN));
  for (int q = 0; q < 3; ++q) {
    for (int k = 0; k < ne; ++k) gw[(size_t)q * ld_w + k] = 0.0;
    for (int s = 0; s < nbs; ++s) gwb[(size_t)q * ld_wb + s] = 0.0;
  }
  for (int i = 0; i < n; ++i) gvol[i] = 0.0;
  if (laminar(c)) {
    TRY(visc_friction_reverse(c, L, true));
    TRY(alloc_vmesh(L));
    DevParams P = c->P;
    LevelDev<double> V = L.view<double>();
    Vlin S(L);
    double *dgw = L.vmesh, *dgwb = L.vmesh + 3 * (size_t)std::max(1, ne);
    const size_t ns = 3 * ((size_t)std::max(1, ne) + (size_t)std::max(1, nbs));
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk((long long)ns)), dim3(NTHR), 0, c->stream, L.vmesh, 0.0, ns);
    const double *nul = nullptr;
    if (ne > 0)
      hipLaunchKernelGGL(k_visc_mesh_side, dim3(nblk(ne)), dim3(NTHR), 0, c->stream, V, P, nul,
                         (const ad::Dual *)S.taud, (const double *)S.gbar, nul, nul, dgw);
    if (nbs > 0)
      hipLaunchKernelGGL(k_visc_mesh_bside, dim3(nblk(nbs)), dim3(NTHR), 0, c->stream, V, P, nul,
                         (const ad::Dual *)S.taud, (const double *)S.gbar, nul, nul, dgwb);
    TRY(check_launch());
    std::vector<double> hw(3 * (size_t)ne), hb(3 * (size_t)nbs), hm(5 * (size_t)n), ha(4 * (size_t)n);
    if (ne > 0) HIPTRY(hipMemcpyAsync(hw.data(), dgw, hw.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (nbs > 0) HIPTRY(hipMemcpyAsync(hb.data(), dgwb, hb.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(hm.data(), S.mbar, hm.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(ha.data(), S.gaf, ha.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < ne; ++k)
      for (int q = 0; q < 3; ++q) gw[(size_t)q * ld_w + L.side_host[k]] += hw[3 * (size_t)k + q];
    for (int s = 0; s < nbs; ++s)
      for (int q = 0; q < 3; ++q) gwb[(size_t)q * ld_wb + s] += hb[3 * (size_t)s + q];
    // the friction's area and normal shares (gaf, gnf: device numbering) and
    // the stress law's normal shares join the pressure term's
    for (int hh = 0; hh < n; ++hh) {
      const int d = L.h2d[hh];
      gvol[hh] += hm[5 * (size_t)d];
      gA[hh] += ha[d];
      for (int q = 0; q < 3; ++q) gN[3 * (size_t)hh + q] += ha[n + 3 * (size_t)d + q] + hm[5 * (size_t)d + 1 + q];
    }
  }
  wb_chain(c->staged[0], n, nbs, gA, gN, gwb, ld_wb);
  return 0;
}

int flite_set_coarse_precision(flite_ctx *c, int fp32) {
  TRY(check_ctx(c));
  if (fp32 != 0 && fp32 != 1) return fail(FLITE_E_ARG, "fp32 must be 0 or 1");
  if (c->nl < 2 || c->lev[1]->fp32 == (fp32 == 1)) return 0;
  TRY(sync(c));
  if (c->lev_alt.empty()) {
    c->lev_alt.resize(c->nl);
    for (int l = 1; l < c->nl; ++l) TRY(build_level(c, l + 1, c->staged[l], c->orders[l], fp32 == 1, c->lev_alt[l]));
  }
  drop_cycle_graph(c);
  for (int l = 1; l < c->nl; ++l) std::swap(c->lev[l], c->lev_alt[l]);
  c->hp.coarse_fp32 = fp32;
  for (auto &q : c->hp_b) q.coarse_fp32 = fp32;
  return 0;
}

// Per level, the lanes per node (1, 4 or 16) of the gather and stage
// kernels that run fastest on this context's current state: each candidate
// runs the level's lane-dependent kernels (pressure and records, time steps
// and viscous/SA passes, gradient, residual gather) `repeats` times after one
// untimed run, timed with HIP events on the context's stream.
int flite_autotune(flite_ctx *c, int repeats, int *lanes_out) {
  TRY(check_ctx(c));
  if (repeats < 1 || !lanes_out) return fail(FLITE_E_ARG, "repeats must be >= 1 and lanes_out non-null");
  TRY(sync(c));
  TRY(harvest_timing(c));
  std::string timed;
  std::swap(timed, c->timed);  // autotuning launches are not counted by kernel timing
  hipEvent_t a = nullptr, b = nullptr;
  int rc = 0;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) rc = fail(FLITE_E_HIP, "hipEventCreate");
  const bool visc = c->hp.reynolds_number > 0.0;
  auto run = [&](Level &L) -> int {
    TRY(op_prep(c, L, true));
    if (visc) TRY(op_step_terms(c, L, true, c->P.sa));
    else TRY(op_timesteps(c, L));
    TRY(op_grad(c, L));
    StageArgs S{};
    S.calc_turb = 1;
    return op_stage(c, L, MODE_RHS, S);
  };
  for (int l = 0; l < c->nl && rc == 0; ++l) {
    Level &L = *c->lev[l];
    int best = L.lanes;
    float best_ms = 1e30f;
    // the timed kernels overwrite the level's derived fields, some of which
    // a cycle keeps from an earlier smoothing step: save and restore them
    const size_t es = L.fp32 ? 4 : 8, nb = (size_t)L.n * L.nbatch;
    const std::pair<void *, size_t> kept[] = {
        {L.p, nb * es},        {L.rec0, 8 * nb * es},  {L.dt, nb * es},         {L.dtt, nb * es},
        {L.mu, nb * es},       {L.vort, nb * es},      {L.divg, nb * es},       {L.tdiff, nb * es},
        {L.diss, 5 * nb * es}, {L.tau, VREC * nb * es}, {L.ws, 3 * nb * es},     {L.rhs, 6 * nb * es},
        {L.grec, (size_t)L.gstride * L.nbatch * es}};
    size_t total = 0;
    for (auto &k : kept) total += k.second;
    char *save = nullptr;
    if (hipMalloc(&save, std::max<size_t>(total, 1)) != hipSuccess) {
      rc = fail(FLITE_E_HIP, "hipMalloc (autotune save)");
      break;
    }
    size_t off = 0;
    for (auto &k : kept) {
      if (k.second && hipMemcpyAsync(save + off, k.first, k.second, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        rc = fail(FLITE_E_HIP, "hipMemcpyAsync (autotune save)");
      off += k.second;
    }
    if (rc == 0) rc = op_prep(c, L, true);
    if (rc == 0) rc = op_flux(c, L);
    for (int lanes : {1, 4, 16}) {
      if (rc != 0) break;
      L.lanes = lanes;
      rc = run(L);
      if (rc == 0 && hipEventRecord(a, c->stream) != hipSuccess) rc = fail(FLITE_E_HIP, "hipEventRecord");
      for (int r = 0; r < repeats && rc == 0; ++r) rc = run(L);
      if (rc == 0 && hipEventRecord(b, c->stream) != hipSuccess) rc = fail(FLITE_E_HIP, "hipEventRecord");
      if (rc == 0) rc = sync(c);
      float ms = 0;
      if (rc == 0 && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = fail(FLITE_E_HIP, "hipEventElapsedTime");
      if (rc == 0 && ms < best_ms) {
        best_ms = ms;
        best = lanes;
      }
    }
    L.lanes = best;
    lanes_out[l] = best;
    // the flux kernel's side-block shape (levels that use side blocks)
    if (rc == 0 && L.ne > 0 && !c->flux_per_side && c->P.ftype == 4) {
      int bs = L.fb_shape;
      float bms = 1e30f;
      rc = op_prep(c, L, true);
      if (rc == 0) rc = op_grad(c, L);
      for (int k = 0; k < FB_NSHAPES && rc == 0; ++k) {
        L.fb_shape = k;
        rc = op_flux(c, L);
        if (rc == 0 && hipEventRecord(a, c->stream) != hipSuccess) rc = fail(FLITE_E_HIP, "hipEventRecord");
        for (int r = 0; r < repeats && rc == 0; ++r) rc = op_flux(c, L);
        if (rc == 0 && hipEventRecord(b, c->stream) != hipSuccess) rc = fail(FLITE_E_HIP, "hipEventRecord");
        if (rc == 0) rc = sync(c);
        float ms = 0;
        if (rc == 0 && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = fail(FLITE_E_HIP, "hipEventElapsedTime");
        if (rc == 0 && ms < bms) {
          bms = ms;
          bs = k;
        }
      }
      L.fb_shape = bs;
    }
    off = 0;
    for (auto &k : kept) {
      if (k.second && hipMemcpyAsync(k.first, save + off, k.second, hipMemcpyDeviceToDevice, c->stream) != hipSuccess &&
          rc == 0)
        rc = fail(FLITE_E_HIP, "hipMemcpyAsync (autotune restore)");
      off += k.second;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == 0) rc = fail(FLITE_E_HIP, "autotune restore");
    (void)hipFree(save);
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  std::swap(timed, c->timed);
  drop_cycle_graph(c);
  return rc;
}

int flite_level_lanes(flite_ctx *c, int level, int *lanes) {
  TRY(check_level(c, level));
  if (!lanes) return fail(FLITE_E_ARG, "null output");
  *lanes = c->lev[level - 1]->lanes;
  return 0;
}

int flite_level_side_block(flite_ctx *c, int level, int *sides, int *nodes) {
  TRY(check_level(c, level));
  if (!sides || !nodes) return fail(FLITE_E_ARG, "null output");
  const int k = c->lev[level - 1]->fb_shape;
  *sides = FB_SHAPE_THR[k];
  *nodes = FB_SHAPE_NODES[k];
  return 0;
}

int flite_batch_select(flite_ctx *c, int batch) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  if (batch < 0 || batch >= c->nbatch) return fail(FLITE_E_ARG, "batch index out of range");
  c->sel = batch;
  return 0;
}

int flite_sync(flite_ctx *c) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  return sync(c);
}

int flite_timing_enable(flite_ctx *c, const char *name) {
  if (!c) return fail(FLITE_E_ARG, "null context");
  TRY(harvest_timing(c));
  c->timed = name ? name : "";
  c->timed_ms = 0;
  c->timed_count = 0;
  return 0;
}

int flite_timing_get(flite_ctx *c, double *total_ms, int *launches) {
  if (!c || !total_ms || !launches) return fail(FLITE_E_ARG, "null argument");
  TRY(harvest_timing(c));
  *total_ms = c->timed_ms;
  *launches = c->timed_count;
  return 0;
}

int flite_cycle_evaluations(flite_ctx *c, int *evals_per_level) {
  TRY(check_ctx(c));
  if (!evals_per_level) return fail(FLITE_E_ARG, "null argument");
  for (int l = 0; l < c->nl; ++l) evals_per_level[l] = l < (int)c->evals.size() ? c->evals[l] : 0;
  return 0;
}

int flite_level_sizes(flite_ctx *c, int level, int *n_nodes, int *n_sides, int *n_bsides) {
  TRY(check_level(c, level));
  Level &L = *c->lev[level - 1];
  if (n_nodes) *n_nodes = L.n;
  if (n_sides) *n_sides = L.ne;
  if (n_bsides) *n_bsides = L.nbs;
  return 0;
}

}  // extern "C"
