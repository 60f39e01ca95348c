// C ABI of the MI355X ICON core: plans, batched operators and the solve.
//
// Reference (tomography/ICON_GPU_v1.2.9/sourceCode/ICON-GPU/):
//   plan setup            MainCode.cu:38-109, nufft_gpu_v8.cu:281-556,1298-1391,1739-1771
//   per-slice solve       MainCode.cu:113-135, Iteration2.cu:33-146, alpha_new2.cu:28-114
//   reprojection          MainCode.cu:142, reProject_NFFT.cu:81-98
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "icon_kernels.h"
#include "icon_plan.h"

using namespace icon;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(ICON_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                  __FILE__, __LINE__);                                                        \
  } while (0)

// Twiddle tables exp(-2 pi i q / n), q < n, per (device, n); computed in
// double and rounded to float.  Held for the life of the process.
std::mutex g_tw_mu;
std::map<std::pair<int, int>, float2*> g_tw;

int twiddles(int n, const float2** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, n);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) { *out = it->second; return ICON_OK; }
  // exp(-2 pi i q / n) in double, rounded to float
  std::vector<float2> w(n);
  for (int q = 0; q < n; q++) {
    const double a = -2.0 * M_PI * (double)q / (double)n;
    w[q] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  // layout read by load_twiddles (icon_fft.h): [dir][slot / 2][t][slot % 2]
  const int TH = n / 16;
  std::vector<float2> h((size_t)2 * TH * 32, make_float2(0.f, 0.f));
  for (int d = 0; d < 2; d++)
    for (int t = 0; t < TH; t++) {
      int ns = 1, P = 0;
      while (ns < n) {
        const int rp = n / ns >= 16 ? 16 : n / ns;
        if (ns > 1) {
          const int step = n / (ns * rp);
          for (int q = 0; q < 16 / rp; q++) {
            const int k = (t + q * TH) & (ns - 1);
            for (int r = 1; r < rp; r++) {
              float2 v = w[(size_t)r * k * step];
              if (d) v.y = -v.y;
              const int slot = P * 16 + q * (rp - 1) + r - 1;
              h[(((size_t)d * 16 + slot / 2) * TH + t) * 2 + slot % 2] = v;
            }
          }
          P++;
        }
        ns *= rp;
      }
    }
  float2* d = nullptr;
  HIPCHK(hipMalloc(&d, sizeof(float2) * h.size()));
  HIPCHK(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
  g_tw[key] = d;
  *out = d;
  return ICON_OK;
}

template <class T>
int dalloc(icon_plan* p, T** ptr, size_t count) {
  const size_t bytes = sizeof(T) * (count ? count : 1);
  HIPCHK(hipMalloc((void**)ptr, bytes));
  p->device_bytes += (int64_t)bytes;
  return ICON_OK;
}

template <class T>
int upload(icon_plan* p, T** ptr, const std::vector<T>& h) {
  int rc = dalloc(p, ptr, h.size());
  if (rc) return rc;
  if (!h.empty()) HIPCHK(hipMemcpy(*ptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return ICON_OK;
}

void free_bins(icon_plan* p) {
  if (p->d_ent_off) (void)hipFree(p->d_ent_off);
  if (p->d_tile_order) (void)hipFree(p->d_tile_order);
  p->d_tile_order = nullptr;
  if (p->d_ent) (void)hipFree(p->d_ent);
  if (p->d_nslot) (void)hipFree(p->d_nslot);
  if (p->d_xoff) (void)hipFree(p->d_xoff);
  if (p->d_ent_px) (void)hipFree(p->d_ent_px);
  if (p->d_ent_py) (void)hipFree(p->d_ent_py);
  if (p->d_ent_of_node) (void)hipFree(p->d_ent_of_node);
  if (p->d_fw) (void)hipFree(p->d_fw);
  if (p->d_fent) (void)hipFree(p->d_fent);
  if (p->d_fent_px) (void)hipFree(p->d_fent_px);
  if (p->d_fent_py) (void)hipFree(p->d_fent_py);
  p->d_ent_off = nullptr; p->d_ent = nullptr; p->d_nslot = nullptr; p->d_xoff = nullptr;
  p->d_ent_px = nullptr; p->d_ent_py = nullptr; p->d_ent_of_node = nullptr; p->d_fw = nullptr;
  p->d_fent = nullptr; p->d_fent_px = nullptr; p->d_fent_py = nullptr;
}

int upload_bins(icon_plan* p) {
  int rc;
  if ((rc = upload(p, &p->d_ent_off, p->bins.off))) return rc;
  if ((rc = upload(p, &p->d_tile_order, p->bins.order))) return rc;
  if ((rc = upload(p, &p->d_ent, p->bins.ent))) return rc;
  if ((rc = upload(p, &p->d_nslot, p->bins.nslot))) return rc;
  if ((rc = upload(p, &p->d_xoff, p->bins.xoff))) return rc;
  if ((rc = upload(p, &p->d_ent_px, p->bins.px))) return rc;
  if ((rc = upload(p, &p->d_ent_py, p->bins.py))) return rc;
  if ((rc = upload(p, &p->d_ent_of_node, p->bins.of_node))) return rc;
  if ((rc = upload(p, &p->d_fent, p->bins.fent))) return rc;
  if ((rc = upload(p, &p->d_fent_px, p->bins.fpx))) return rc;
  if ((rc = upload(p, &p->d_fent_py, p->bins.fpy))) return rc;
  // weighted node value per bin entry, per slice
  if ((rc = dalloc(p, &p->d_fw, (size_t)p->S_max * p->bins.ent.size()))) return rc;
  return ICON_OK;
}

void drop_graph(icon_plan* p) {
  for (int i = 0; i < kGraphCache; i++) {
    if (p->graph_exec[i]) (void)hipGraphExecDestroy(p->graph_exec[i]);
    p->graph_exec[i] = nullptr;
    p->graph_key[i] = GraphKey{};
    p->graph_used[i] = 0;
  }
  p->data_sino = nullptr;
  p->data_S = 0;
}

// Eight grid columns per tile: two blocks fit in one CU's LDS at n0 = 1024
// (measured faster than 16 columns at one block per CU).
int default_col_block(int n0) { (void)n0; return 8; }

struct DevGuard {
  int prev = -1;
  explicit DevGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DevGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

int check_launch(hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(ICON_ERR_DEVICE, "%s launch failed: %s", what, hipGetErrorString(e));
  return ICON_OK;
}

// ---------------------------------------------------------------- operators

int op_forward(icon_plan* p, int S, const void* img, int cplx, float2* out, hipStream_t st) {
  const Geometry& g = p->g;
  RowFwdArgs ra{};
  ra.rows = S * g.N0; ra.N0 = g.N0; ra.N1 = g.N1; ra.cplx = cplx; ra.update = 0;
  ra.m = cplx ? nullptr : (float*)img;
  ra.rhat = cplx ? (const float2*)img : nullptr;
  ra.ck = p->d_ck; ra.tw = p->d_tw1; ra.out = p->d_tmp;
  zero_band(p->bins, &ra.zlo, &ra.zhi);
  int rc = check_launch(launch_row_fwd(g.n1, ra, st), "row_fwd");
  if (rc) return rc;
  ColArgs ca{};
  ca.K = p->g.K;
  ca.S = S; ca.N0 = g.N0; ca.N1 = g.N1; ca.M = g.M; ca.n1 = g.n1;
  ca.in = p->d_tmp; ca.tw = p->d_tw0; ca.ent_off = p->d_ent_off; ca.tile_order = p->d_tile_order; ca.split_min = split_min_entries(p->bins); ca.ent = p->d_ent; ca.xoff = p->d_xoff; ca.E = (int)p->bins.ent.size(); ca.ent_px = p->d_ent_px; ca.ent_py = p->d_ent_py;
  ca.fent = p->d_fent; ca.fent_px = p->d_fent_px; ca.fent_py = p->d_fent_py;
  ca.phix = p->d_phix; ca.phiy = p->d_phiy; ca.ix0 = p->d_ix0; ca.part = p->d_part;
  if ((rc = check_launch(launch_col(g.n0, p->bins.C, false, ca, st), "col_fwd"))) return rc;
  CombineArgs cb{};
  cb.S = S; cb.M = g.M; cb.N0 = g.N0; cb.nz = g.nz; cb.mode = 3; cb.part = p->d_part;
  cb.nslot = p->d_nslot; cb.w = p->d_w; cb.fw = out;
  return check_launch(launch_combine(cb, st), "combine");
}

// A^H of the per-entry weighted values already in p->d_fw.
int op_adjoint_ent(icon_plan* p, int S, float2* img, hipStream_t st) {
  const Geometry& g = p->g;
  const float2* f = p->d_fw;
  ColArgs ca{};
  ca.K = p->g.K;
  ca.S = S; ca.N0 = g.N0; ca.N1 = g.N1; ca.M = g.M; ca.n1 = g.n1;
  ca.out = p->d_tmp; ca.tw = p->d_tw0; ca.ent_off = p->d_ent_off; ca.tile_order = p->d_tile_order; ca.split_min = split_min_entries(p->bins); ca.ent = p->d_ent; ca.xoff = p->d_xoff; ca.E = (int)p->bins.ent.size(); ca.ent_px = p->d_ent_px; ca.ent_py = p->d_ent_py;
  ca.fent = p->d_fent; ca.fent_px = p->d_fent_px; ca.fent_py = p->d_fent_py;
  ca.phix = p->d_phix; ca.phiy = p->d_phiy; ca.ix0 = p->d_ix0; ca.fw = f;
  zero_band(p->bins, &ca.zlo, &ca.zhi);
  int rc = check_launch(launch_col(g.n0, p->bins.C, true, ca, st), "col_adj");
  if (rc) return rc;
  RowAdjArgs ra{};
  ra.rows = S * g.N0; ra.N0 = g.N0; ra.N1 = g.N1; ra.mode = 0; ra.in = p->d_tmp;
  ra.ck = p->d_ck; ra.tw = p->d_tw1; ra.img = img;
  zero_band(p->bins, &ra.zlo, &ra.zhi);
  return check_launch(launch_row_adj(g.n1, ra, st), "row_adj");
}

// A^H of node vectors f [S][M].
int op_adjoint(icon_plan* p, int S, const float2* f, float2* img, hipStream_t st) {
  int rc = check_launch(launch_node_to_ent(f, p->d_fw, p->d_ent, S, p->g.M, (int)p->bins.ent.size(), st),
                        "node_to_ent");
  if (rc) return rc;
  return op_adjoint_ent(p, S, img, st);
}

int op_sino(icon_plan* p, int S, const float* sino, float2* b, hipStream_t st) {
  const Geometry& g = p->g;
  SinoArgs sa{};
  sa.K = p->g.K;
  sa.rows = S * g.nz; sa.N0 = g.N0; sa.nz = g.nz; sa.skip = g.cv_skip; sa.sino = sino; sa.ck1 = p->d_ck1;
  sa.phi1 = p->d_phi1;
  sa.i10 = p->d_i10; sa.tw = p->d_tw0; sa.b = b;
  return check_launch(launch_sino(g.n0, sa, st), "sino");
}

// The whole solve of S slices, enqueued on st (captured into a graph by the caller
// when requested).
int enqueue_solve(icon_plan* p, int S, const float* sino, float* m, const float* m_init, float* reproj,
                  double* resid, double* alpha_tr, const icon_solve_opts& o, hipStream_t st) {
  const Geometry& g = p->g;
  const int iters = o.pre + o.icon + o.post;
  int rc;
  if (!o.reuse_data) {
    // b = NFFT1D of every projection row; fw = w b; bnorm = sum w |b|^2
    if ((rc = op_sino(p, S, sino, p->d_b, st))) return rc;
    CombineArgs c0{};
    c0.S = S; c0.M = g.M; c0.N0 = g.N0; c0.nz = g.nz; c0.mode = 2; c0.b = p->d_b; c0.w = p->d_w;
    c0.fw = p->d_fw; c0.bnorm = p->d_bnorm; c0.nslot = p->d_nslot; c0.part = p->d_part;
    c0.E = (int)p->bins.ent.size(); c0.ent_of_node = p->d_ent_of_node;
    if ((rc = check_launch(launch_combine(c0, st), "combine_b"))) return rc;
    // A^H W b
    if ((rc = op_adjoint_ent(p, S, p->d_ahwb, st))) return rc;
  }
  // m = 0 (Iteration2.cu:33) or the warm start
  const size_t NN = (size_t)g.N0 * g.N1;
  if (m_init) rc = check_launch(launch_copy(m, m_init, (size_t)S * NN, st), "copy");
  else rc = check_launch(launch_zero(m, (size_t)S * NN, st), "zero");
  if (rc) return rc;

  RowFwdArgs rf{};
  rf.rows = S * g.N0; rf.N0 = g.N0; rf.N1 = g.N1; rf.data_type = o.data_type; rf.thr = o.threshold;
  rf.m = m; rf.grad = p->d_grad; rf.alpha = p->d_alpha; rf.ck = p->d_ck; rf.tw = p->d_tw1; rf.out = p->d_tmp;
  ColArgs ca{};
  ca.K = p->g.K;
  ca.S = S; ca.N0 = g.N0; ca.N1 = g.N1; ca.M = g.M; ca.n1 = g.n1; ca.in = p->d_tmp; ca.out = p->d_tmp;
  ca.tw = p->d_tw0; ca.ent_off = p->d_ent_off; ca.tile_order = p->d_tile_order; ca.split_min = split_min_entries(p->bins); ca.ent = p->d_ent; ca.xoff = p->d_xoff; ca.E = (int)p->bins.ent.size(); ca.ent_px = p->d_ent_px; ca.ent_py = p->d_ent_py;
  ca.fent = p->d_fent; ca.fent_px = p->d_fent_px; ca.fent_py = p->d_fent_py; ca.phix = p->d_phix; ca.phiy = p->d_phiy;
  ca.ix0 = p->d_ix0; ca.part = p->d_part; ca.fw = p->d_fw;
  CombineArgs cm{};
  cm.S = S; cm.M = g.M; cm.N0 = g.N0; cm.nz = g.nz; cm.mode = 0; cm.part = p->d_part; cm.nslot = p->d_nslot;
  cm.w = p->d_w; cm.b = p->d_b; cm.fw = p->d_fw; cm.bnorm = p->d_bnorm; cm.resid = resid;
  cm.resid_stride = iters + 1; cm.iomit = g.iomit;
  cm.E = (int)p->bins.ent.size(); cm.ent_of_node = p->d_ent_of_node;
  CombineArgs cr = cm;
  cr.mode = 1; cr.resid = nullptr; cr.rowpart = p->d_rowpart; cr.alpha = p->d_alpha; cr.alpha_trace = alpha_tr;
  cr.alpha_stride = iters;
  RowAdjArgs ra{};
  ra.rows = S * g.N0; ra.N0 = g.N0; ra.N1 = g.N1; ra.mode = 1; ra.in = p->d_tmp; ra.ck = p->d_ck;
  ra.tw = p->d_tw1; ra.ahwb = p->d_ahwb; ra.grad = p->d_grad; ra.rowpart = p->d_rowpart;
  ra.fwd_out = p->d_tmp;   // the row pass of A rhat fused in
  zero_band(p->bins, &rf.zlo, &rf.zhi);
  zero_band(p->bins, &ca.zlo, &ca.zhi);
  zero_band(p->bins, &ra.zlo, &ra.zhi);
  auto phase = [&](int it) { return (o.pre <= it && it < iters - o.post) ? 2 : 1; };

  for (int it = 0; it < iters; it++) {
    // f = A m (with the previous iteration's update fused in front)
    rf.cplx = 0;
    rf.update = it == 0 ? 0 : phase(it - 1);
    if ((rc = check_launch(launch_row_fwd(g.n1, rf, st), "row_fwd"))) return rc;
    if ((rc = check_launch(launch_col(g.n0, p->bins.C, false, ca, st), "col_fwd"))) return rc;
    // R(m_it), fw = w f
    cm.resid_index = it;
    cm.frep = nullptr;
    if ((rc = check_launch(launch_combine(cm, st), "combine_m"))) return rc;
    // rhat = A^H fw - A^H W b, row sums of |rhat|^2, and the row pass of A rhat
    if ((rc = check_launch(launch_col(g.n0, p->bins.C, true, ca, st), "col_adj"))) return rc;
    if ((rc = check_launch(launch_row_adj(g.n1, ra, st), "row_adj"))) return rc;
    // alpha = |rhat|^2 / sum w |A rhat|^2
    if ((rc = check_launch(launch_col(g.n0, p->bins.C, false, ca, st), "col_fwd_c"))) return rc;
    cr.alpha_index = it;
    if ((rc = check_launch(launch_combine(cr, st), "combine_r"))) return rc;
  }
  // last update, then A m_final for the final residual and the reprojection
  if (iters > 0 || resid || reproj) {
    rf.cplx = 0;
    rf.update = iters > 0 ? phase(iters - 1) : 0;
    if ((rc = check_launch(launch_row_fwd(g.n1, rf, st), "row_fwd_final"))) return rc;
  }
  if (resid || reproj) {
    if ((rc = check_launch(launch_col(g.n0, p->bins.C, false, ca, st), "col_fwd_final"))) return rc;
    cm.resid_index = iters;
    cm.frep = reproj ? p->d_frep : nullptr;
    if ((rc = check_launch(launch_combine(cm, st), "combine_final"))) return rc;
    if (reproj) {
      ReprojArgs rp{};
      rp.K = p->g.K;
      rp.S = S; rp.N0 = g.N0; rp.sigma = g.sig0; rp.frep = p->d_frep; rp.ck1 = p->d_ck1;
      rp.phi1 = p->d_phi1; rp.i10 = p->d_i10; rp.tw = p->d_tw0; rp.out = reproj;
      if ((rc = check_launch(launch_reproj(g.n0, rp, st), "reproj"))) return rc;
    }
  }
  return ICON_OK;
}

int prepare_col(int n0, int n1, int C, int K) {
  // sets the LDS attribute outside any capture
  if (!col_tile_ok(n0, n1, C, K))
    return fail(ICON_ERR_ARG, "column tile %d unsupported for grid %d x %d with %d window taps", C, n0, n1, K);
  hipError_t e = set_col_attrs(n0, C, K);
  if (e != hipSuccess) return fail(ICON_ERR_DEVICE, "hipFuncSetAttribute: %s", hipGetErrorString(e));
  return ICON_OK;
}

}  // namespace

void icon::set_last_error(const char* msg) { g_err = msg; }

// ====================================================================== ABI
extern "C" {

const char* icon_last_error(void) { return g_err.c_str(); }

const char* icon_version(void) { return "icon-mi355x 0.1 (gfx950)"; }

int icon_device_count(int* count) {
  if (!count) return fail(ICON_ERR_ARG, "icon_device_count: null pointer");
  HIPCHK(hipGetDeviceCount(count));
  return ICON_OK;
}

int icon_device_compute_units(int device, int* cus) {
  if (!cus) return fail(ICON_ERR_ARG, "icon_device_compute_units: null pointer");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  *cus = prop.multiProcessorCount;
  return ICON_OK;
}

int icon_set_device(int device) {
  HIPCHK(hipSetDevice(device));
  return ICON_OK;
}

int icon_get_device(int* device) {
  if (!device) return fail(ICON_ERR_ARG, "icon_get_device: null pointer");
  HIPCHK(hipGetDevice(device));
  return ICON_OK;
}

int icon_stream_create(void** stream) {
  if (!stream) return fail(ICON_ERR_ARG, "icon_stream_create: null pointer");
  hipStream_t st = nullptr;
  HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *stream = st;
  return ICON_OK;
}

int icon_stream_destroy(void* stream) {
  if (stream) HIPCHK(hipStreamDestroy((hipStream_t)stream));
  return ICON_OK;
}

int icon_stream_synchronize(void* stream) {
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return ICON_OK;
}

int icon_event_create(void** event) {
  if (!event) return fail(ICON_ERR_ARG, "icon_event_create: null pointer");
  hipEvent_t ev = nullptr;
  HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  *event = ev;
  return ICON_OK;
}

int icon_event_destroy(void* event) {
  if (event) HIPCHK(hipEventDestroy((hipEvent_t)event));
  return ICON_OK;
}

int icon_event_record(void* event, void* stream) {
  if (!event) return fail(ICON_ERR_ARG, "icon_event_record: null event");
  HIPCHK(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return ICON_OK;
}

int icon_stream_wait_event(void* stream, void* event) {
  if (!event) return fail(ICON_ERR_ARG, "icon_stream_wait_event: null event");
  HIPCHK(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
  return ICON_OK;
}

int icon_malloc(void** dptr, size_t bytes) {
  if (!dptr) return fail(ICON_ERR_ARG, "icon_malloc: null pointer");
  HIPCHK(hipMalloc(dptr, bytes ? bytes : 1));
  return ICON_OK;
}

int icon_free(void* dptr) {
  if (dptr) HIPCHK(hipFree(dptr));
  return ICON_OK;
}

int icon_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return ICON_OK;
}

int icon_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return ICON_OK;
}

int icon_memcpy_d2d_async(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return ICON_OK;
  if (!dst || !src) return fail(ICON_ERR_ARG, "icon_memcpy_d2d_async: null pointer");
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return ICON_OK;
}

int icon_memset_d(void* dst, int value, size_t bytes) {
  HIPCHK(hipMemset(dst, value, bytes));
  return ICON_OK;
}

int icon_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return ICON_OK;
}

int icon_geometry_tables(int nx, int thickness, int nz, const float* tilt_deg, icon_plan_info* info,
                         float* nodes, float* weights, float* phix, float* phiy, int* ix0, int* iy0, float* ck,
                         float* ck1, float* phi1, int* i10) {
  if (!tilt_deg) return fail(ICON_ERR_ARG, "icon_geometry_tables: null tilt angles");
  Geometry g;
  std::string err;
  if (!build_geometry(g, nx, thickness > 0 ? thickness : nx, nz, tilt_deg, &err))
    return fail(ICON_ERR_ARG, "%s", err.c_str());
  auto put = [](auto* dst, const auto& v) {
    if (dst) std::memcpy(dst, v.data(), sizeof(v[0]) * v.size());
  };
  put(nodes, g.nodes); put(weights, g.w); put(phix, g.phix); put(phiy, g.phiy); put(ix0, g.ix0);
  put(iy0, g.iy0); put(ck, g.ck); put(ck1, g.ck1); put(phi1, g.phi1); put(i10, g.i10);
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->N0 = g.N0; info->N1 = g.N1; info->nz = g.nz; info->M = g.M; info->n0 = g.n0; info->n1 = g.n1;
    info->m = g.K == kTaps ? kWindowM : g.K / 2; info->window = g.window; info->iomit = g.iomit; info->sigma0 = g.sig0; info->sigma1 = g.sig1;
    info->b0 = g.b0; info->b1 = g.b1;
  }
  return ICON_OK;
}

namespace {
int plan_create(icon_plan** out, int device, int nx, int thickness, int nz, const float* tilt_deg, int max_slices,
                bool cv, int window = ICON_WINDOW_GAUSS);
}

int icon_plan_create(icon_plan** out, int device, int nx, int thickness, int nz, const float* tilt_deg,
                     int max_slices) {
  return plan_create(out, device, nx, thickness, nz, tilt_deg, max_slices, false);
}

int icon_plan_create_cv(icon_plan** out, int device, int nx, int thickness, int nz, const float* tilt_deg,
                        int max_slices) {
  return plan_create(out, device, nx, thickness, nz, tilt_deg, max_slices, true);
}

int icon_plan_create_window(icon_plan** out, int device, int nx, int thickness, int nz, const float* tilt_deg,
                            int max_slices, int cross_validation, int window) {
  return plan_create(out, device, nx, thickness, nz, tilt_deg, max_slices, cross_validation != 0, window);
}

}  // extern "C"

namespace {
int plan_create(icon_plan** out, int device, int nx, int thickness, int nz, const float* tilt_deg, int max_slices,
                bool cv, int window) {
  if (!out || !tilt_deg) return fail(ICON_ERR_ARG, "icon_plan_create: null pointer");
  *out = nullptr;
  if (max_slices < 1) return fail(ICON_ERR_ARG, "icon_plan_create: max_slices must be >= 1");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(ICON_ERR_ARG, "icon_plan_create: no device %d", device);
  DevGuard dg(device);
  HIPCHK(hipSetDevice(device));
  auto* p = new icon_plan();
  p->device = device;
  std::string err;
  if (!build_geometry(p->g, nx, thickness > 0 ? thickness : nx, nz, tilt_deg, &err, cv, window)) {
    delete p;
    return fail(ICON_ERR_ARG, "%s", err.c_str());
  }
  p->S_max = max_slices;
  build_bins(p->g, default_col_block(p->g.n0), p->bins);
  const Geometry& g = p->g;
  int rc = ICON_OK;
  auto bail = [&](int code) {
    icon_plan_destroy(p);
    return code;
  };
  if ((rc = upload(p, &p->d_ck, g.ck))) return bail(rc);
  if ((rc = upload(p, &p->d_ck1, g.ck1))) return bail(rc);
  if ((rc = upload(p, &p->d_w, g.w))) return bail(rc);
  if ((rc = upload(p, &p->d_phix, g.phix))) return bail(rc);
  if ((rc = upload(p, &p->d_phiy, g.phiy))) return bail(rc);
  if ((rc = upload(p, &p->d_ix0, g.ix0))) return bail(rc);
  if ((rc = upload(p, &p->d_phi1, g.phi1))) return bail(rc);
  if ((rc = upload(p, &p->d_i10, g.i10))) return bail(rc);
  if ((rc = upload_bins(p))) return bail(rc);
  if ((rc = twiddles(g.n0, &p->d_tw0))) return bail(rc);
  if ((rc = twiddles(g.n1, &p->d_tw1))) return bail(rc);
  const size_t S = (size_t)max_slices, M = (size_t)g.M, NN = (size_t)g.N0 * g.N1;
  if ((rc = dalloc(p, &p->d_b, S * M))) return bail(rc);
  if ((rc = dalloc(p, &p->d_part, S * M * kMaxSlot))) return bail(rc);
  if ((rc = dalloc(p, &p->d_ahwb, S * NN))) return bail(rc);
  if ((rc = dalloc(p, &p->d_grad, S * NN))) return bail(rc);
  if ((rc = dalloc(p, &p->d_tmp, S * g.N0 * (size_t)g.n1))) return bail(rc);
  if ((rc = dalloc(p, &p->d_frep, S * g.N0))) return bail(rc);
  if ((rc = dalloc(p, &p->d_rowpart, S * g.N0))) return bail(rc);
  if ((rc = dalloc(p, &p->d_bnorm, S))) return bail(rc);
  if ((rc = dalloc(p, &p->d_alpha, S))) return bail(rc);
  // alpha starts at 0 so an update before any step is a no-op
  {
    hipError_t e = hipMemset(p->d_alpha, 0, sizeof(double) * S);
    if (e != hipSuccess) return bail(fail(ICON_ERR_DEVICE, "hipMemset: %s", hipGetErrorString(e)));
  }
  {
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming);
    if (e != hipSuccess) return bail(fail(ICON_ERR_DEVICE, "stream/event create: %s", hipGetErrorString(e)));
  }
  if ((rc = prepare_col(g.n0, g.n1, p->bins.C, g.K))) return bail(rc);
  *out = p;
  return ICON_OK;
}
}  // namespace

extern "C" {

int icon_plan_destroy(icon_plan* p) {
  if (!p) return ICON_OK;
  DevGuard dg(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  drop_graph(p);
  void* ptrs[] = {p->d_ck, p->d_ck1, p->d_w, p->d_phix, p->d_phiy, p->d_ix0, p->d_phi1, p->d_i10,
                  p->d_b, p->d_part, p->d_ahwb, p->d_grad, p->d_tmp, p->d_frep, p->d_rowpart,
                  p->d_bnorm, p->d_alpha};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  free_bins(p);
  if (p->ev_in) (void)hipEventDestroy(p->ev_in);
  if (p->ev_out) (void)hipEventDestroy(p->ev_out);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return ICON_OK;
}

int icon_plan_get_info(const icon_plan* p, icon_plan_info* info) {
  if (!p || !info) return fail(ICON_ERR_ARG, "icon_plan_get_info: null pointer");
  const Geometry& g = p->g;
  info->N0 = g.N0; info->N1 = g.N1; info->nz = g.nz; info->M = g.M; info->n0 = g.n0; info->n1 = g.n1;
  info->m = g.K == kTaps ? kWindowM : g.K / 2; info->window = g.window; info->iomit = g.iomit; info->sigma0 = g.sig0; info->sigma1 = g.sig1;
  info->b0 = g.b0; info->b1 = g.b1; info->max_slices = p->S_max; info->col_block = p->bins.C;
  info->device_bytes = p->device_bytes;
  return ICON_OK;
}

int icon_plan_set_col_block(icon_plan* p, int C) {
  if (!p) return fail(ICON_ERR_ARG, "icon_plan_set_col_block: null plan");
  if (!col_tile_ok(p->g.n0, p->g.n1, C, p->g.K))
    return fail(ICON_ERR_ARG, "column tile %d unsupported for grid %d x %d", C, p->g.n0, p->g.n1);
  DevGuard dg(p->device);
  HIPCHK(hipStreamSynchronize(p->stream));
  drop_graph(p);
  free_bins(p);
  build_bins(p->g, C, p->bins);
  int rc = upload_bins(p);
  if (rc) return rc;
  return prepare_col(p->g.n0, p->g.n1, C, p->g.K);
}

int icon_nufft_forward_real(icon_plan* p, int S, const void* d_img, void* d_f, void* stream) {
  if (!p || !d_img || !d_f) return fail(ICON_ERR_ARG, "icon_nufft_forward_real: null pointer");
  if (S < 1 || S > p->S_max) return fail(ICON_ERR_CAPACITY, "batch %d outside 1..%d", S, p->S_max);
  DevGuard dg(p->device);
  return op_forward(p, S, d_img, 0, (float2*)d_f, (hipStream_t)stream);
}

int icon_nufft_forward_complex(icon_plan* p, int S, const void* d_img, void* d_f, void* stream) {
  if (!p || !d_img || !d_f) return fail(ICON_ERR_ARG, "icon_nufft_forward_complex: null pointer");
  if (S < 1 || S > p->S_max) return fail(ICON_ERR_CAPACITY, "batch %d outside 1..%d", S, p->S_max);
  DevGuard dg(p->device);
  return op_forward(p, S, d_img, 1, (float2*)d_f, (hipStream_t)stream);
}

int icon_nufft_adjoint(icon_plan* p, int S, const void* d_f, void* d_img, void* stream) {
  if (!p || !d_img || !d_f) return fail(ICON_ERR_ARG, "icon_nufft_adjoint: null pointer");
  if (S < 1 || S > p->S_max) return fail(ICON_ERR_CAPACITY, "batch %d outside 1..%d", S, p->S_max);
  DevGuard dg(p->device);
  return op_adjoint(p, S, (const float2*)d_f, (float2*)d_img, (hipStream_t)stream);
}

int icon_sino_to_b(icon_plan* p, int S, const void* d_sino, void* d_b, void* stream) {
  if (!p || !d_sino || !d_b) return fail(ICON_ERR_ARG, "icon_sino_to_b: null pointer");
  if (S < 1) return fail(ICON_ERR_ARG, "batch %d < 1", S);
  DevGuard dg(p->device);
  return op_sino(p, S, (const float*)d_sino, (float2*)d_b, (hipStream_t)stream);
}

int icon_fft_rows(int n, int rows, int sign, void* d_data, void* stream) {
  if (!d_data || rows < 0 || (sign != 1 && sign != -1)) return fail(ICON_ERR_ARG, "icon_fft_rows: bad argument");
  if (n != 128 && n != 256 && n != 512 && n != 1024 && n != 2048)
    return fail(ICON_ERR_ARG, "icon_fft_rows: n=%d unsupported", n);
  if (rows == 0) return ICON_OK;
  const float2* tw = nullptr;
  int rc = twiddles(n, &tw);
  if (rc) return rc;
  return check_launch(launch_fft_rows(n, (float2*)d_data, rows, sign, tw, (hipStream_t)stream), "fft_rows");
}

int icon_fft_2d(int n0, int n1, int count, int sign, void* d_data, void* stream) {
  if (!d_data || count < 0 || (sign != 1 && sign != -1)) return fail(ICON_ERR_ARG, "icon_fft_2d: bad argument");
  for (int n : {n0, n1})
    if (n != 128 && n != 256 && n != 512 && n != 1024 && n != 2048)
      return fail(ICON_ERR_ARG, "icon_fft_2d: n=%d unsupported", n);
  if (count == 0) return ICON_OK;
  const float2 *tw0 = nullptr, *tw1 = nullptr;
  int rc = twiddles(n0, &tw0);
  if (!rc) rc = twiddles(n1, &tw1);
  if (rc) return rc;
  const int C = default_col_block(n0);
  if ((rc = check_launch(launch_fft_rows(n1, (float2*)d_data, count * n0, sign, tw1, (hipStream_t)stream), "fft_rows")))
    return rc;
  return check_launch(launch_fft_cols(n0, C, (float2*)d_data, n1, count, sign, tw0, (hipStream_t)stream), "fft_cols");
}

int icon_reconstruct(icon_plan* p, int S, const void* d_sino, void* d_m, const void* d_m_init, void* d_reproj,
                     void* d_resid, void* d_alpha, const icon_solve_opts* opts, void* stream) {
  if (!p || !d_sino || !d_m || !opts) return fail(ICON_ERR_ARG, "icon_reconstruct: null pointer");
  if (S < 1 || S > p->S_max) return fail(ICON_ERR_CAPACITY, "batch %d outside 1..%d", S, p->S_max);
  const icon_solve_opts& o = *opts;
  if (o.pre < 0 || o.icon < 0 || o.post < 0) return fail(ICON_ERR_ARG, "negative iteration count");
  if (o.data_type != 1 && o.data_type != 2) return fail(ICON_ERR_ARG, "data_type must be 1 or 2");
  if (o.reuse_data && !(p->data_sino == d_sino && p->data_S == S))
    return fail(ICON_ERR_ARG, "reuse_data: the previous solve on this plan had other sinograms or batch");
  DevGuard dg(p->device);
  hipStream_t user = (hipStream_t)stream;
  HIPCHK(hipEventRecord(p->ev_in, user));
  HIPCHK(hipStreamWaitEvent(p->stream, p->ev_in, 0));
  int rc = ICON_OK;
  if (o.use_graph) {
    GraphKey key;
    key.S = S; key.sino = d_sino; key.m = d_m; key.m_init = d_m_init; key.reproj = d_reproj;
    key.resid = d_resid; key.alpha = d_alpha; key.pre = o.pre; key.icon = o.icon; key.post = o.post;
    key.data_type = o.data_type; key.reuse = o.reuse_data ? 1 : 0; key.thr = o.threshold;
    int slot = -1;
    for (int i = 0; i < kGraphCache; i++)
      if (p->graph_exec[i] && p->graph_key[i] == key) slot = i;
    if (slot < 0) {
      slot = 0;   // least recently used (empty slots have time 0)
      for (int i = 1; i < kGraphCache; i++)
        if (p->graph_used[i] < p->graph_used[slot]) slot = i;
      if (p->graph_exec[slot]) (void)hipGraphExecDestroy(p->graph_exec[slot]);
      p->graph_exec[slot] = nullptr;
      p->graph_key[slot] = GraphKey{};
      HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
      rc = enqueue_solve(p, S, (const float*)d_sino, (float*)d_m, (const float*)d_m_init, (float*)d_reproj,
                         (double*)d_resid, (double*)d_alpha, o, p->stream);
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamEndCapture(p->stream, &graph);
      if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
      }
      if (e != hipSuccess) return fail(ICON_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&p->graph_exec[slot], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        p->graph_exec[slot] = nullptr;
        return fail(ICON_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
      }
      p->graph_key[slot] = key;
    }
    p->graph_used[slot] = ++p->graph_clock;
    HIPCHK(hipGraphLaunch(p->graph_exec[slot], p->stream));
  } else {
    rc = enqueue_solve(p, S, (const float*)d_sino, (float*)d_m, (const float*)d_m_init, (float*)d_reproj,
                       (double*)d_resid, (double*)d_alpha, o, p->stream);
    if (rc) return rc;
  }
  p->data_sino = d_sino;
  p->data_S = S;
  HIPCHK(hipEventRecord(p->ev_out, p->stream));
  HIPCHK(hipStreamWaitEvent(user, p->ev_out, 0));
  return rc;
}

int icon_reconstruct_host(icon_plan* p, int S, const float* sino, float* m, const float* m_init, float* reproj,
                          double* resid, double* alpha, const icon_solve_opts* opts) {
  if (!p || !sino || !m || !opts) return fail(ICON_ERR_ARG, "icon_reconstruct_host: null pointer");
  if (S < 1 || S > p->S_max) return fail(ICON_ERR_CAPACITY, "batch %d outside 1..%d", S, p->S_max);
  if (opts->reuse_data) return fail(ICON_ERR_ARG, "icon_reconstruct_host: reuse_data needs device buffers");
  DevGuard dg(p->device);
  const Geometry& g = p->g;
  const int iters = opts->pre + opts->icon + opts->post;
  const size_t nsino = (size_t)S * g.M, nimg = (size_t)S * g.N0 * g.N1;
  void *ds = nullptr, *dm = nullptr, *dmi = nullptr, *drp = nullptr, *drs = nullptr, *dal = nullptr;
  int rc = ICON_OK;
  auto cleanup = [&]() {
    for (void* q : {ds, dm, dmi, drp, drs, dal})
      if (q) (void)hipFree(q);
  };
#define TRY(expr)                                              \
  do {                                                         \
    hipError_t e_ = (expr);                                    \
    if (e_ != hipSuccess) {                                    \
      cleanup();                                               \
      return fail(ICON_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    }                                                          \
  } while (0)
  TRY(hipMalloc(&ds, nsino * sizeof(float)));
  TRY(hipMalloc(&dm, nimg * sizeof(float)));
  if (m_init) TRY(hipMalloc(&dmi, nimg * sizeof(float)));
  if (reproj) TRY(hipMalloc(&drp, (size_t)S * g.N0 * sizeof(float)));
  if (resid) TRY(hipMalloc(&drs, (size_t)S * (iters + 1) * sizeof(double)));
  if (alpha && iters > 0) TRY(hipMalloc(&dal, (size_t)S * iters * sizeof(double)));
  TRY(hipMemcpy(ds, sino, nsino * sizeof(float), hipMemcpyHostToDevice));
  if (m_init) TRY(hipMemcpy(dmi, m_init, nimg * sizeof(float), hipMemcpyHostToDevice));
  rc = icon_reconstruct(p, S, ds, dm, dmi, drp, drs, dal, opts, nullptr);
  if (rc) { cleanup(); return rc; }
  TRY(hipDeviceSynchronize());
  TRY(hipMemcpy(m, dm, nimg * sizeof(float), hipMemcpyDeviceToHost));
  if (reproj) TRY(hipMemcpy(reproj, drp, (size_t)S * g.N0 * sizeof(float), hipMemcpyDeviceToHost));
  if (resid) TRY(hipMemcpy(resid, drs, (size_t)S * (iters + 1) * sizeof(double), hipMemcpyDeviceToHost));
  if (alpha && iters > 0) TRY(hipMemcpy(alpha, dal, (size_t)S * iters * sizeof(double), hipMemcpyDeviceToHost));
#undef TRY
  cleanup();
  return ICON_OK;
}

int icon_time_kernel(icon_plan* p, int kernel, int S, int reps, double* avg_ms) {
  if (!p || !avg_ms || reps < 1) return fail(ICON_ERR_ARG, "icon_time_kernel: bad argument");
  if (S < 1 || S > p->S_max) return fail(ICON_ERR_CAPACITY, "batch %d outside 1..%d", S, p->S_max);
  DevGuard dg(p->device);
  const Geometry& g = p->g;
  hipStream_t st = p->stream;
  p->data_sino = nullptr;   // the timed kernels overwrite the solve's data buffers
  RowFwdArgs rf{};
  // the row pass with the ICON update fused, as in every iteration after the
  // first; the image it updates lives in the (then stale) A^H W b buffer
  rf.rows = S * g.N0; rf.N0 = g.N0; rf.N1 = g.N1; rf.cplx = 0; rf.update = 2; rf.data_type = 1;
  rf.m = reinterpret_cast<float*>(p->d_ahwb); rf.grad = p->d_grad; rf.alpha = p->d_alpha;
  rf.ck = p->d_ck; rf.tw = p->d_tw1; rf.out = p->d_tmp;
  ColArgs ca{};
  ca.K = p->g.K;
  ca.S = S; ca.N0 = g.N0; ca.N1 = g.N1; ca.M = g.M; ca.n1 = g.n1; ca.in = p->d_tmp; ca.out = p->d_tmp;
  ca.tw = p->d_tw0; ca.ent_off = p->d_ent_off; ca.tile_order = p->d_tile_order; ca.split_min = split_min_entries(p->bins); ca.ent = p->d_ent; ca.xoff = p->d_xoff; ca.E = (int)p->bins.ent.size(); ca.ent_px = p->d_ent_px; ca.ent_py = p->d_ent_py;
  ca.fent = p->d_fent; ca.fent_px = p->d_fent_px; ca.fent_py = p->d_fent_py; ca.phix = p->d_phix; ca.phiy = p->d_phiy;
  ca.ix0 = p->d_ix0; ca.part = p->d_part; ca.fw = p->d_fw;
  CombineArgs cm{};
  cm.S = S; cm.M = g.M; cm.N0 = g.N0; cm.nz = g.nz; cm.mode = 0; cm.part = p->d_part; cm.nslot = p->d_nslot;
  cm.w = p->d_w; cm.b = p->d_b; cm.fw = p->d_fw; cm.bnorm = p->d_bnorm;
  cm.E = (int)p->bins.ent.size(); cm.ent_of_node = p->d_ent_of_node;
  RowAdjArgs ra{};
  ra.rows = S * g.N0; ra.N0 = g.N0; ra.N1 = g.N1; ra.mode = 1; ra.in = p->d_tmp; ra.ck = p->d_ck;
  ra.tw = p->d_tw1; ra.ahwb = p->d_ahwb; ra.grad = p->d_grad; ra.rowpart = p->d_rowpart;
  ra.fwd_out = p->d_tmp;
  zero_band(p->bins, &rf.zlo, &rf.zhi);
  zero_band(p->bins, &ca.zlo, &ca.zhi);
  zero_band(p->bins, &ra.zlo, &ra.zhi);
  // the step-size combine (mode 1) after the combine of A m, as in a solve
  CombineArgs cr = cm;
  cr.mode = 1; cr.rowpart = p->d_rowpart; cr.alpha = p->d_alpha;
  SinoArgs sa{};
  sa.K = p->g.K;
  sa.rows = S * g.nz; sa.N0 = g.N0; sa.sino = reinterpret_cast<const float*>(p->d_b); sa.ck1 = p->d_ck1;
  sa.phi1 = p->d_phi1; sa.i10 = p->d_i10; sa.tw = p->d_tw0; sa.b = p->d_fw;
  auto launch = [&](int k) -> hipError_t {
    switch (k) {
      case ICON_K_ROW_FWD: return launch_row_fwd(g.n1, rf, st);
      case ICON_K_COL_FWD: return launch_col(g.n0, p->bins.C, false, ca, st);
      case ICON_K_COMBINE: return launch_combine(cm, st);
      case ICON_K_COL_ADJ: return launch_col(g.n0, p->bins.C, true, ca, st);
      case ICON_K_ROW_ADJ: return launch_row_adj(g.n1, ra, st);
      case ICON_K_SINO: return launch_sino(g.n0, sa, st);
      case ICON_K_COMBINE_ALPHA: return launch_combine(cr, st);
      default: return hipErrorInvalidValue;
    }
  };
  if (kernel < ICON_K_ROW_FWD || kernel > ICON_K_COMBINE_ALPHA)
    return fail(ICON_ERR_ARG, "icon_time_kernel: unknown kernel %d", kernel);
  // The launch sequence of one iteration, as the solve enqueues it; the
  // timed kernel runs inside it, so that each timed launch finds its input as
  // in a solve (just written by its predecessor, partly in the caches) and the
  // device busy (an idle gap before a launch lowers the clock it starts at).
  // The sinogram transform runs once per solve and is timed back to back.
  const int iter_seq[7] = {ICON_K_ROW_FWD, ICON_K_COL_FWD, ICON_K_COMBINE, ICON_K_COL_ADJ,
                           ICON_K_ROW_ADJ, ICON_K_COL_FWD, ICON_K_COMBINE_ALPHA};
  const int sino_seq[1] = {ICON_K_SINO};
  const int* seq = kernel == ICON_K_SINO ? sino_seq : iter_seq;
  const int nseq = kernel == ICON_K_SINO ? 1 : 7;
  // Each timed launch carries events that take the kernel's own start and end
  // (hipExtLaunchKernel), not those of extra marker packets around it; all
  // launches are enqueued before one synchronisation.
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  hipError_t e = hipSuccess;
  for (size_t i = 0; i < ev.size() && e == hipSuccess; i++) e = hipEventCreate(&ev[i]);
  for (int i = 0; i < nseq && e == hipSuccess; i++) e = launch(seq[i]);  // untimed warm-up pass
  for (int r = 0; r < reps && e == hipSuccess;) {
    for (int i = 0; i < nseq && e == hipSuccess; i++) {
      const bool timed = seq[i] == kernel && r < reps;
      if (timed) set_launch_events(ev[2 * r], ev[2 * r + 1]);
      e = launch(seq[i]);
      if (timed) {
        set_launch_events(nullptr, nullptr);
        r++;
      }
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  std::vector<float> t(reps, 0.f);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = hipEventElapsedTime(&t[r], ev[2 * r], ev[2 * r + 1]);
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (e != hipSuccess) return fail(ICON_ERR_DEVICE, "icon_time_kernel: %s", hipGetErrorString(e));
  // median launch duration: robust to a launch slowed by other work on the device
  std::sort(t.begin(), t.end());
  *avg_ms = reps % 2 ? t[reps / 2] : 0.5 * ((double)t[reps / 2 - 1] + t[reps / 2]);
  return ICON_OK;
}

int icon_plan_autotune(icon_plan* p, int S, int* col_block, double* ms_per_width) {
  if (!p || !col_block) return fail(ICON_ERR_ARG, "icon_plan_autotune: null pointer");
  const int widths[2] = {8, 16};
  double best = 0.0;
  int best_c = 0;
  for (int i = 0; i < 2; i++) {
    const int C = widths[i];
    if (ms_per_width) ms_per_width[i] = 0.0;
    if (!col_tile_ok(p->g.n0, p->g.n1, C, p->g.K)) continue;
    int rc = icon_plan_set_col_block(p, C);
    double tf = 0.0, ta = 0.0;
    if (!rc) rc = icon_time_kernel(p, ICON_K_COL_FWD, S, 3, &tf);
    if (!rc) rc = icon_time_kernel(p, ICON_K_COL_ADJ, S, 3, &ta);
    if (rc) return rc;
    const double t = 2.0 * tf + ta;
    if (ms_per_width) ms_per_width[i] = t;
    if (best_c == 0 || t < best) { best = t; best_c = C; }
  }
  if (best_c == 0) return fail(ICON_ERR_ARG, "icon_plan_autotune: no supported column tile");
  if (p->bins.C != best_c) {
    const int rc = icon_plan_set_col_block(p, best_c);
    if (rc) return rc;
  }
  *col_block = best_c;
  return ICON_OK;
}

int icon_copy_probe(void* d_dst, const void* d_src, size_t bytes, int reps, double* avg_ms) {
  if (!d_dst || !d_src || !avg_ms || reps < 1 || bytes % 16 != 0)
    return fail(ICON_ERR_ARG, "icon_copy_probe: bad argument (bytes must be a multiple of 16)");
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  hipError_t e = launch_copy4((float4*)d_dst, (const float4*)d_src, bytes / 16, nullptr);
  if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy4((float4*)d_dst, (const float4*)d_src, bytes / 16, nullptr);
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(ICON_ERR_DEVICE, "icon_copy_probe: %s", hipGetErrorString(e));
  *avg_ms = (double)ms / reps;
  return ICON_OK;
}

}  // extern "C"
