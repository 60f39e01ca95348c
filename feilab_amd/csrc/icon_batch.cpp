// Packed solve of several tilt series of uneven width, thickness, tilt count
// and slice count in the same launches.
//
// The reference solves one slice at a time per device and derives the grid,
// window and node set per series (MainCode.cu:63-71, the slice loop
// MainCode.cu:113-162).  Here every series keeps its own plan (geometry
// tables, bins, per-slice buffers), and a batch enqueues each kernel of the
// iteration once per size class instead of once per series: the row kernels
// once per row length n1, the column kernels once per (n0, tile width,
// window), the node combine once for all series.  Each packed launch reads
// the argument block of a series from a device table and the (series, block)
// pair of each of its workgroups from an item table built at batch creation.
// The whole solve is captured into one HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "icon_kernels.h"
#include "icon_plan.h"

using namespace icon;

namespace {

int bfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  set_last_error(buf);
  return code;
}

#define BCHK(expr)                                                                               \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return bfail(ICON_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                   __LINE__);                                                                    \
  } while (0)

struct Items {
  int key0 = 0, key1 = 0, key2 = 0;  // size class (meaning per kernel)
  std::vector<int2> h;
  int2* d = nullptr;
  int n() const { return (int)h.size(); }
};

struct SolveKey {
  std::vector<const void*> ptrs;
  int pre = -1, icon = -1, post = -1, data_type = 0;
  float thr = 0;
  bool operator==(const SolveKey& o) const {
    return ptrs == o.ptrs && pre == o.pre && icon == o.icon && post == o.post && data_type == o.data_type &&
           thr == o.thr;
  }
};

}  // namespace

struct icon_batch {
  int device = 0;
  std::vector<icon_plan*> plans;
  std::vector<int> S;
  std::vector<int> C_built;   // column tile of each plan when the items were built
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // item tables per size class
  std::vector<Items> row_fwd, row_adj, col, sino, reproj;
  // the spread's items: those of col, by their tile's entry count, largest
  // first across the class's series (each tile's blocks stay together)
  std::vector<Items> col_adj;
  Items combine;
  // argument blocks, one per series (device copies written before each solve)
  std::vector<RowFwdArgs> h_rf;
  std::vector<RowAdjArgs> h_ra;
  std::vector<ColArgs> h_ca;
  std::vector<CombineArgs> h_cm;
  std::vector<SinoArgs> h_sa;
  std::vector<ReprojArgs> h_rp;
  RowFwdArgs* d_rf = nullptr;
  RowAdjArgs* d_ra = nullptr;
  ColArgs* d_ca = nullptr;
  CombineArgs* d_cm = nullptr;
  SinoArgs* d_sa = nullptr;
  ReprojArgs* d_rp = nullptr;
  hipGraphExec_t graph = nullptr;
  SolveKey graph_key;
};

namespace {

void free_items(std::vector<Items>& v) {
  for (auto& it : v)
    if (it.d) (void)hipFree(it.d);
  v.clear();
}

int upload_items(Items& it) {
  if (it.d) (void)hipFree(it.d);
  it.d = nullptr;
  if (it.h.empty()) return ICON_OK;
  BCHK(hipMalloc(&it.d, sizeof(int2) * it.h.size()));
  BCHK(hipMemcpy(it.d, it.h.data(), sizeof(int2) * it.h.size(), hipMemcpyHostToDevice));
  return ICON_OK;
}

// (series, block) items of every size class; called at creation and after the
// column tiles change.
int build_items(icon_batch* b) {
  free_items(b->row_fwd);
  free_items(b->row_adj);
  free_items(b->col);
  free_items(b->col_adj);
  free_items(b->sino);
  free_items(b->reproj);
  if (b->combine.d) (void)hipFree(b->combine.d);
  b->combine = Items{};
  const int ns = (int)b->plans.size();
  b->C_built.assign(ns, 0);
  for (int i = 0; i < ns; i++) b->C_built[i] = b->plans[i]->bins.C;
  auto group = [](std::vector<Items>& v, int k0, int k1, int k2) -> Items& {
    for (auto& it : v)
      if (it.key0 == k0 && it.key1 == k1 && it.key2 == k2) return it;
    v.push_back(Items{});
    v.back().key0 = k0; v.back().key1 = k1; v.back().key2 = k2;
    return v.back();
  };
  for (int i = 0; i < ns; i++) {
    const icon_plan* p = b->plans[i];
    const Geometry& g = p->g;
    const int S = b->S[i];
    const int rows = S * g.N0;
    Items& rf = group(b->row_fwd, g.n1, 0, 0);
    const int rpf = row_fwd_rows_per_block(g.n1);
    for (int k = 0; k < (rows + rpf - 1) / rpf; k++) rf.h.push_back(make_int2(i, k));
    Items& ra = group(b->row_adj, g.n1, 0, 0);
    const int rpa = row_adj_rows_per_block(g.n1);
    for (int k = 0; k < (rows + rpa - 1) / rpa; k++) ra.h.push_back(make_int2(i, k));
    if (!col_tile_ok(g.n0, g.n1, p->bins.C, g.K))
      return bfail(ICON_ERR_ARG, "series %d: column tile %d unsupported for grid %d x %d", i, p->bins.C, g.n0, g.n1);
    Items& cl = group(b->col, g.n0, p->bins.C, g.K);
    for (int k = 0; k < S * (g.n1 / p->bins.C); k++) cl.h.push_back(make_int2(i, k));
    Items& sn = group(b->sino, g.n0, 0, 0);
    const int rps = sino_rows_per_block(g.n0);
    for (int k = 0; k < (S * g.nz + rps - 1) / rps; k++) sn.h.push_back(make_int2(i, k));
    Items& rp = group(b->reproj, g.n0, 0, 0);
    for (int k = 0; k < S; k++) rp.h.push_back(make_int2(i, k));
    for (int k = 0; k < S; k++) b->combine.h.push_back(make_int2(i, k));
  }
  // the spread's order: a stable sort of each class's items by the entry
  // count of the tile the block runs (the plan's own largest-first tile order
  // maps a series' block k to tile bins.order[k / S])
  for (const Items& c : b->col) {
    Items a = c;
    a.d = nullptr;
    auto cnt = [&](const int2& q) {
      const icon_plan* p = b->plans[q.x];
      const int t = p->bins.order[q.y / b->S[q.x]];
      return p->bins.off[t + 1] - p->bins.off[t];
    };
    std::stable_sort(a.h.begin(), a.h.end(), [&](const int2& u, const int2& v) { return cnt(u) > cnt(v); });
    b->col_adj.push_back(std::move(a));
  }
  int rc;
  for (auto* v : {&b->row_fwd, &b->row_adj, &b->col, &b->col_adj, &b->sino, &b->reproj})
    for (auto& it : *v)
      if ((rc = upload_items(it))) return rc;
  if ((rc = upload_items(b->combine))) return rc;
  for (auto& it : b->col) {
    hipError_t e = set_col_multi_attrs(it.key0, it.key1, it.key2);
    if (e != hipSuccess) return bfail(ICON_ERR_DEVICE, "column kernel attributes: %s", hipGetErrorString(e));
  }
  return ICON_OK;
}

void drop_batch_graph(icon_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_key = SolveKey{};
}

// Argument blocks of every series for a solve on the given buffers.
void fill_args(icon_batch* b, const float* const* sino, float* const* m, float* const* reproj, double* const* resid,
               double* const* alpha, int iters) {
  const int ns = (int)b->plans.size();
  b->h_rf.assign(ns, RowFwdArgs{});
  b->h_ra.assign(ns, RowAdjArgs{});
  b->h_ca.assign(ns, ColArgs{});
  b->h_cm.assign(ns, CombineArgs{});
  b->h_sa.assign(ns, SinoArgs{});
  b->h_rp.assign(ns, ReprojArgs{});
  for (int i = 0; i < ns; i++) {
    icon_plan* p = b->plans[i];
    const Geometry& g = p->g;
    const int S = b->S[i];
    const int E = (int)p->bins.ent.size();
    RowFwdArgs& rf = b->h_rf[i];
    rf.rows = S * g.N0; rf.N0 = g.N0; rf.N1 = g.N1; rf.m = m[i]; rf.grad = p->d_grad; rf.alpha = p->d_alpha;
    rf.ck = p->d_ck; rf.tw = p->d_tw1; rf.out = p->d_tmp;
    RowAdjArgs& ra = b->h_ra[i];
    ra.rows = S * g.N0; ra.N0 = g.N0; ra.N1 = g.N1; ra.in = p->d_tmp; ra.ck = p->d_ck; ra.tw = p->d_tw1;
    ra.img = p->d_ahwb; ra.ahwb = p->d_ahwb; ra.grad = p->d_grad; ra.rowpart = p->d_rowpart; ra.fwd_out = p->d_tmp;
    ColArgs& ca = b->h_ca[i];
    ca.S = S; ca.N0 = g.N0; ca.N1 = g.N1; ca.M = g.M; ca.n1 = g.n1; ca.E = E; ca.K = g.K;
    ca.in = p->d_tmp; ca.out = p->d_tmp; ca.tw = p->d_tw0; ca.ent_off = p->d_ent_off; ca.tile_order = p->d_tile_order; ca.split_min = split_min_entries(p->bins); ca.ent = p->d_ent;
    ca.xoff = p->d_xoff; ca.phix = p->d_phix; ca.phiy = p->d_phiy; ca.ix0 = p->d_ix0; ca.ent_px = p->d_ent_px;
    ca.ent_py = p->d_ent_py; ca.fent = p->d_fent; ca.fent_px = p->d_fent_px; ca.fent_py = p->d_fent_py;
    ca.part = p->d_part; ca.fw = p->d_fw;
    zero_band(p->bins, &rf.zlo, &rf.zhi);
    zero_band(p->bins, &ra.zlo, &ra.zhi);
    zero_band(p->bins, &ca.zlo, &ca.zhi);
    CombineArgs& cm = b->h_cm[i];
    cm.S = S; cm.M = g.M; cm.N0 = g.N0; cm.nz = g.nz; cm.E = E; cm.part = p->d_part; cm.nslot = p->d_nslot;
    cm.w = p->d_w; cm.b = p->d_b; cm.fw = p->d_fw; cm.ent_of_node = p->d_ent_of_node; cm.bnorm = p->d_bnorm;
    cm.resid = resid ? resid[i] : nullptr; cm.resid_stride = iters + 1; cm.rowpart = p->d_rowpart;
    cm.alpha = p->d_alpha; cm.alpha_trace = alpha ? alpha[i] : nullptr; cm.alpha_stride = iters;
    cm.frep = p->d_frep; cm.iomit = g.iomit;
    SinoArgs& sa = b->h_sa[i];
    sa.rows = S * g.nz; sa.N0 = g.N0; sa.K = g.K; sa.nz = g.nz; sa.skip = g.cv_skip; sa.sino = sino[i];
    sa.ck1 = p->d_ck1; sa.phi1 = p->d_phi1; sa.i10 = p->d_i10; sa.tw = p->d_tw0; sa.b = p->d_b;
    ReprojArgs& rp = b->h_rp[i];
    rp.S = S; rp.N0 = g.N0; rp.K = g.K; rp.sigma = g.sig0; rp.frep = p->d_frep; rp.ck1 = p->d_ck1;
    rp.phi1 = p->d_phi1; rp.i10 = p->d_i10; rp.tw = p->d_tw0; rp.out = reproj ? reproj[i] : nullptr;
  }
}

template <class T>
int upload_args(T* d, const std::vector<T>& h, hipStream_t st) {
  BCHK(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
  return ICON_OK;
}

int ck(hipError_t e, const char* what) {
  if (e != hipSuccess) return bfail(ICON_ERR_DEVICE, "%s launch failed: %s", what, hipGetErrorString(e));
  return ICON_OK;
}

// The whole packed solve, enqueued on st (the order of enqueue_solve in the
// single-series path, each step once per size class).
int enqueue_batch(icon_batch* b, float* const* m, const float* const* m_init, bool any_resid, bool any_reproj,
                  const icon_solve_opts& o, hipStream_t st) {
  const int ns = (int)b->plans.size();
  const int iters = o.pre + o.icon + o.post;
  int rc;
  auto col_all = [&](bool adj) -> int {
    for (auto& it : adj ? b->col_adj : b->col)
      if ((rc = ck(launch_col_multi(it.key0, it.key1, it.key2, adj, b->d_ca, it.d, it.n(), st),
                   adj ? "col_adj" : "col_fwd")))
        return rc;
    return ICON_OK;
  };
  auto row_fwd_all = [&](int update) -> int {
    RowFwdLaunch L{0, update, o.data_type, o.threshold};
    for (auto& it : b->row_fwd)
      if ((rc = ck(launch_row_fwd_multi(it.key0, b->d_rf, it.d, it.n(), L, st), "row_fwd"))) return rc;
    return ICON_OK;
  };
  auto row_adj_all = [&](int mode, bool fwd) -> int {
    for (auto& it : b->row_adj)
      if ((rc = ck(launch_row_adj_multi(it.key0, b->d_ra, it.d, it.n(), mode, fwd, st), "row_adj"))) return rc;
    return ICON_OK;
  };
  auto combine = [&](int mode, int ri, int ai, int with_frep) -> int {
    CombineLaunch L{mode, ri, ai, 1, with_frep};
    return ck(launch_combine_multi(b->d_cm, b->combine.d, b->combine.n(), L, st), "combine");
  };
  // data set-up: b, w b, bnorm, A^H W b
  for (auto& it : b->sino)
    if ((rc = ck(launch_sino_multi(it.key0, b->d_sa, it.d, it.n(), st), "sino"))) return rc;
  if ((rc = combine(2, 0, 0, 0))) return rc;
  if ((rc = col_all(true))) return rc;
  if ((rc = row_adj_all(0, false))) return rc;
  // m = 0 or the warm start
  for (int i = 0; i < ns; i++) {
    const size_t n = (size_t)b->S[i] * b->plans[i]->g.N0 * b->plans[i]->g.N1;
    hipError_t e = m_init && m_init[i] ? launch_copy(m[i], m_init[i], n, st) : launch_zero(m[i], n, st);
    if ((rc = ck(e, "init"))) return rc;
  }
  auto phase = [&](int it) { return (o.pre <= it && it < iters - o.post) ? 2 : 1; };
  for (int it = 0; it < iters; it++) {
    if ((rc = row_fwd_all(it == 0 ? 0 : phase(it - 1)))) return rc;
    if ((rc = col_all(false))) return rc;
    if ((rc = combine(0, it, 0, 0))) return rc;
    if ((rc = col_all(true))) return rc;
    if ((rc = row_adj_all(1, true))) return rc;
    if ((rc = col_all(false))) return rc;
    if ((rc = combine(1, 0, it, 0))) return rc;
  }
  if (iters > 0 || any_resid || any_reproj)
    if ((rc = row_fwd_all(iters > 0 ? phase(iters - 1) : 0))) return rc;
  if (any_resid || any_reproj) {
    if ((rc = col_all(false))) return rc;
    if ((rc = combine(0, iters, 0, any_reproj ? 1 : 0))) return rc;
    if (any_reproj)
      for (auto& it : b->reproj)
        if ((rc = ck(launch_reproj_multi(it.key0, b->d_rp, it.d, it.n(), st), "reproj"))) return rc;
  }
  return ICON_OK;
}

}  // namespace

extern "C" {

int icon_batch_destroy(icon_batch* b) {
  if (!b) return ICON_OK;
  int prev = -1;
  (void)hipGetDevice(&prev);
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_batch_graph(b);
  free_items(b->row_fwd);
  free_items(b->row_adj);
  free_items(b->col);
  free_items(b->col_adj);
  free_items(b->sino);
  free_items(b->reproj);
  if (b->combine.d) (void)hipFree(b->combine.d);
  for (void* q : {(void*)b->d_rf, (void*)b->d_ra, (void*)b->d_ca, (void*)b->d_cm, (void*)b->d_sa, (void*)b->d_rp})
    if (q) (void)hipFree(q);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  if (prev >= 0) (void)hipSetDevice(prev);
  return ICON_OK;
}

int icon_batch_create(icon_batch** out, int nseries, icon_plan* const* plans, const int* slices) {
  if (!out || !plans || !slices || nseries < 1) return bfail(ICON_ERR_ARG, "icon_batch_create: bad argument");
  *out = nullptr;
  for (int i = 0; i < nseries; i++) {
    if (!plans[i]) return bfail(ICON_ERR_ARG, "icon_batch_create: series %d has no plan", i);
    if (plans[i]->device != plans[0]->device)
      return bfail(ICON_ERR_ARG, "icon_batch_create: series %d is on device %d, series 0 on %d", i,
                   plans[i]->device, plans[0]->device);
    if (slices[i] < 1 || slices[i] > plans[i]->S_max)
      return bfail(ICON_ERR_CAPACITY, "icon_batch_create: series %d: %d slices outside 1..%d", i, slices[i],
                   plans[i]->S_max);
    for (int j = 0; j < i; j++)
      if (plans[j] == plans[i]) return bfail(ICON_ERR_ARG, "icon_batch_create: series %d and %d share a plan", j, i);
  }
  auto* b = new icon_batch();
  b->device = plans[0]->device;
  b->plans.assign(plans, plans + nseries);
  b->S.assign(slices, slices + nseries);
  int prev = -1;
  (void)hipGetDevice(&prev);
  auto done = [&](int rc) {
    if (rc) icon_batch_destroy(b);
    else *out = b;
    if (prev >= 0) (void)hipSetDevice(prev);
    return rc;
  };
  if (hipSetDevice(b->device) != hipSuccess) return done(bfail(ICON_ERR_DEVICE, "icon_batch_create: hipSetDevice"));
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&b->d_rf, sizeof(RowFwdArgs) * nseries);
  if (e == hipSuccess) e = hipMalloc(&b->d_ra, sizeof(RowAdjArgs) * nseries);
  if (e == hipSuccess) e = hipMalloc(&b->d_ca, sizeof(ColArgs) * nseries);
  if (e == hipSuccess) e = hipMalloc(&b->d_cm, sizeof(CombineArgs) * nseries);
  if (e == hipSuccess) e = hipMalloc(&b->d_sa, sizeof(SinoArgs) * nseries);
  if (e == hipSuccess) e = hipMalloc(&b->d_rp, sizeof(ReprojArgs) * nseries);
  if (e != hipSuccess) return done(bfail(ICON_ERR_DEVICE, "icon_batch_create: %s", hipGetErrorString(e)));
  return done(build_items(b));
}

int icon_batch_reconstruct(icon_batch* b, const void* const* d_sino, void* const* d_m, const void* const* d_m_init,
                           void* const* d_reproj, void* const* d_resid, void* const* d_alpha,
                           const icon_solve_opts* opts, void* stream) {
  if (!b || !d_sino || !d_m || !opts) return bfail(ICON_ERR_ARG, "icon_batch_reconstruct: null pointer");
  const icon_solve_opts& o = *opts;
  if (o.pre < 0 || o.icon < 0 || o.post < 0) return bfail(ICON_ERR_ARG, "negative iteration count");
  if (o.data_type != 1 && o.data_type != 2) return bfail(ICON_ERR_ARG, "data_type must be 1 or 2");
  if (o.reuse_data) return bfail(ICON_ERR_ARG, "icon_batch_reconstruct: reuse_data is not supported for batches");
  const int ns = (int)b->plans.size();
  for (int i = 0; i < ns; i++)
    if (!d_sino[i] || !d_m[i]) return bfail(ICON_ERR_ARG, "icon_batch_reconstruct: series %d has no sinograms or image", i);
  const int iters = o.pre + o.icon + o.post;
  bool any_resid = false, any_reproj = false;
  SolveKey key;
  key.pre = o.pre; key.icon = o.icon; key.post = o.post; key.data_type = o.data_type; key.thr = o.threshold;
  for (int i = 0; i < ns; i++) {
    any_resid |= d_resid && d_resid[i];
    any_reproj |= d_reproj && d_reproj[i];
    key.ptrs.push_back(d_sino[i]);
    key.ptrs.push_back(d_m[i]);
    key.ptrs.push_back(d_m_init ? d_m_init[i] : nullptr);
    key.ptrs.push_back(d_reproj ? d_reproj[i] : nullptr);
    key.ptrs.push_back(d_resid ? d_resid[i] : nullptr);
    key.ptrs.push_back(d_alpha ? d_alpha[i] : nullptr);
  }
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != b->device) BCHK(hipSetDevice(b->device));
  int rc = ICON_OK;
  hipStream_t user = (hipStream_t)stream;
  std::vector<float*> m(ns);
  std::vector<const float*> sino(ns), mi(ns);
  std::vector<float*> rp(ns);
  std::vector<double*> rs(ns), al(ns);
  for (int i = 0; i < ns; i++) {
    sino[i] = (const float*)d_sino[i];
    m[i] = (float*)d_m[i];
    mi[i] = d_m_init ? (const float*)d_m_init[i] : nullptr;
    rp[i] = d_reproj ? (float*)d_reproj[i] : nullptr;
    rs[i] = d_resid ? (double*)d_resid[i] : nullptr;
    al[i] = d_alpha ? (double*)d_alpha[i] : nullptr;
  }
  // the plans' own buffers are overwritten: their data reuse is void
  for (auto* p : b->plans) {
    p->data_sino = nullptr;
    p->data_S = 0;
  }
  // a plan whose column tile changed since the items were built (its bins
  // have another tile count): rebuild the items before any launch
  for (int i = 0; i < ns && rc == ICON_OK; i++)
    if (b->plans[i]->bins.C != b->C_built[i]) {
      drop_batch_graph(b);
      rc = build_items(b);
    }
  do {
    if (rc) break;
    hipError_t e = hipEventRecord(b->ev_in, user);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->ev_in, 0);
    if (e != hipSuccess) { rc = bfail(ICON_ERR_DEVICE, "icon_batch_reconstruct: %s", hipGetErrorString(e)); break; }
    if (!(b->graph && b->graph_key == key)) {
      drop_batch_graph(b);
      fill_args(b, sino.data(), m.data(), rp.data(), rs.data(), al.data(), iters);
      if ((rc = upload_args(b->d_rf, b->h_rf, b->stream)) || (rc = upload_args(b->d_ra, b->h_ra, b->stream)) ||
          (rc = upload_args(b->d_ca, b->h_ca, b->stream)) || (rc = upload_args(b->d_cm, b->h_cm, b->stream)) ||
          (rc = upload_args(b->d_sa, b->h_sa, b->stream)) || (rc = upload_args(b->d_rp, b->h_rp, b->stream)))
        break;
      e = hipStreamSynchronize(b->stream);
      if (e == hipSuccess) e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) { rc = bfail(ICON_ERR_DEVICE, "capture: %s", hipGetErrorString(e)); break; }
      rc = enqueue_batch(b, m.data(), mi.data(), any_resid, any_reproj, o, b->stream);
      hipGraph_t graph = nullptr;
      e = hipStreamEndCapture(b->stream, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        break;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        b->graph = nullptr;
        rc = bfail(ICON_ERR_DEVICE, "graph: %s", hipGetErrorString(e));
        break;
      }
      b->graph_key = key;
    }
    e = hipGraphLaunch(b->graph, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->ev_out, b->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(user, b->ev_out, 0);
    if (e != hipSuccess) rc = bfail(ICON_ERR_DEVICE, "icon_batch_reconstruct: %s", hipGetErrorString(e));
  } while (0);
  if (prev >= 0 && prev != b->device) (void)hipSetDevice(prev);
  return rc;
}

int icon_batch_launches_per_iteration(const icon_batch* b, int* launches) {
  if (!b || !launches) return bfail(ICON_ERR_ARG, "icon_batch_launches_per_iteration: null pointer");
  *launches = (int)b->row_fwd.size() + (int)b->row_adj.size() + 3 * (int)b->col.size() + 2;
  return ICON_OK;
}

// Column-tile autotuner of a batch: for every column size class (n0, window)
// each tile width the class supports is timed on the class's packed
// interpolation and spread launches (2 + 1, as in an iteration) over the
// plans' current buffers, and the fastest is kept for all plans of the class.
int icon_batch_autotune(icon_batch* b, int* col_blocks) {
  if (!b) return bfail(ICON_ERR_ARG, "icon_batch_autotune: null batch");
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != b->device) BCHK(hipSetDevice(b->device));
  drop_batch_graph(b);
  const int ns = (int)b->plans.size();
  // classes by (n0, K)
  std::map<std::pair<int, int>, std::vector<int>> cls;
  for (int i = 0; i < ns; i++) cls[{b->plans[i]->g.n0, b->plans[i]->g.K}].push_back(i);
  int rc = ICON_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  std::vector<const float*> sino(ns, nullptr);
  std::vector<float*> m(ns, nullptr), rp(ns, nullptr);
  std::vector<double*> rs(ns, nullptr), al(ns, nullptr);
  for (auto& kv : cls) {
    const int n0 = kv.first.first, K = kv.first.second;
    double best = -1.0;
    int best_c = 0;
    for (int C : {8, 16}) {
      bool ok = true;
      for (int i : kv.second) ok = ok && col_tile_ok(n0, b->plans[i]->g.n1, C, K);
      if (!ok) continue;
      for (int i : kv.second)
        if (b->plans[i]->bins.C != C && (rc = icon_plan_set_col_block(b->plans[i], C)))
          return bfail(rc, "icon_batch_autotune: %s", icon_last_error());
      if ((rc = build_items(b))) return rc;
      fill_args(b, sino.data(), m.data(), rp.data(), rs.data(), al.data(), 0);
      if ((rc = upload_args(b->d_ca, b->h_ca, b->stream))) return rc;
      const Items *it = nullptr, *ia = nullptr;
      for (auto& c : b->col)
        if (c.key0 == n0 && c.key1 == C && c.key2 == K) it = &c;
      for (auto& c : b->col_adj)
        if (c.key0 == n0 && c.key1 == C && c.key2 == K) ia = &c;
      if (!it || !ia) continue;
      double t = 0.0;
      for (int rep = 0; rep < 4 && rc == ICON_OK; rep++) {
        hipError_t e = hipEventRecord(e0, b->stream);
        for (int k = 0; k < 3 && e == hipSuccess; k++)
          e = k == 1 ? launch_col_multi(n0, C, K, true, b->d_ca, ia->d, ia->n(), b->stream)
                     : launch_col_multi(n0, C, K, false, b->d_ca, it->d, it->n(), b->stream);
        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) rc = bfail(ICON_ERR_DEVICE, "icon_batch_autotune: %s", hipGetErrorString(e));
        if (rep > 0) t += ms;   // the first pass is a warm-up
      }
      if (rc) break;
      if (best < 0 || t < best) { best = t; best_c = C; }
    }
    if (rc) break;
    for (int i : kv.second)
      if (best_c && b->plans[i]->bins.C != best_c && (rc = icon_plan_set_col_block(b->plans[i], best_c)))
        return bfail(rc, "icon_batch_autotune: %s", icon_last_error());
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = build_items(b);
  if (!rc && col_blocks)
    for (int i = 0; i < ns; i++) col_blocks[i] = b->plans[i]->bins.C;
  if (prev >= 0 && prev != b->device) (void)hipSetDevice(prev);
  return rc;
}

}  // extern "C"
