// Host orchestration behind the C ABI: items -> images -> features -> pairing + Pfaffian ->
// projection, and the local-energy pipeline.  No allocation: every buffer is carved from the
// caller's workspace.

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "nh_common.h"
#include "nh_internal.h"

namespace nh {

// ---- measurement hook: HIP events around every launch of a kernel class, on its stream
struct ProfRec {
  hipEvent_t a, b;
  int kid;
};
static std::mutex g_prof_mu;
static std::vector<ProfRec> g_prof;
static std::atomic<bool> g_prof_on{false};

// Brackets the launches of one kernel class with two events on the launch stream while the
// profiling hook is enabled; a record whose events could not be created or recorded is dropped.
struct ProfScope {
  hipStream_t st;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(hipStream_t s, int k) : st(s), kid(k) {
    if (!g_prof_on.load(std::memory_order_relaxed)) return;
    if (hipEventCreate(&a) != hipSuccess) {
      a = nullptr;
      return;
    }
    if (hipEventCreate(&b) != hipSuccess || hipEventRecord(a, st) != hipSuccess) drop();
  }
  ~ProfScope() {
    if (!a || !b) return;
    if (hipEventRecord(b, st) != hipSuccess) {
      drop();
      return;
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back({a, b, kid});
  }
  void drop() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    a = b = nullptr;
  }
};

int check_model(const nh_model* m) {
  if (!m) return NH_ERR_ARG;
  if (m->kind != NH_KIND_PFAFFIAN && m->kind != NH_KIND_BCS) return NH_ERR_ARG;
  if (m->LX < 3 || m->LY < 3 || m->LX * m->LY > 1024) return NH_ERR_SHAPE;
  if (m->n_layers < 0 || m->n_up < 0 || m->n_dn < 0) return NH_ERR_ARG;
  if (m->network == NH_NET_RESCONV || m->network == NH_NET_RESCONV_HEAD) {
    if (!resconv_supported(m)) return NH_ERR_SHAPE;
    if (m->precision == NH_PREC_MIXED && !m->conv_w_lp) return NH_ERR_UNSUPPORTED;
  } else if (m->network != NH_NET_CONV) {
    return NH_ERR_ARG;
  } else if (m->channels != 8 && m->channels != 16 && m->channels != 32) {
    return NH_ERR_SHAPE;
  }
  const int N = m->LX * m->LY;
  if (m->n_up > N || m->n_dn > N) return NH_ERR_SHAPE;
  if (!m->conv_w || !m->conv_b || !m->wt) return NH_ERR_ARG;
  if (m->kind == NH_KIND_PFAFFIAN && ((m->n_up + m->n_dn) % 2 || m->n_up + m->n_dn > 1024)) return NH_ERR_SHAPE;
  if (m->kind == NH_KIND_BCS && (m->n_up != m->n_dn || m->n_up > 512)) return NH_ERR_SHAPE;
  if (m->precision != NH_PREC_FP32 && m->precision != NH_PREC_MIXED) return NH_ERR_ARG;
  if (m->network == NH_NET_CONV && m->precision == NH_PREC_MIXED && (!m->conv_w_lp || !cnn_mfma_supported(m)))
    return NH_ERR_UNSUPPORTED;
  return NH_OK;
}

static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* p) : base((char*)p) {}
  template <class T>
  T* take(size_t n) {
    T* p = base ? (T*)(base + off) : nullptr;
    off += align_up(n * sizeof(T));
    return p;
  }
};

static int ipi_of(const nh_model* m) { return m->project ? kGroup : 1; }

struct ItemBuffers {
  float* Y;
  float* img_sign;
  double* img_logabs;
  float* gtri;  // Pfaffian triangles above n = 64 / BCS matrices above n = 192, one per image of a chunk
  float* net;   // residual-CNN activations of a chunk on lattices outside its tile kernels (or NULL)
  int64_t chunk;
};

static ItemBuffers carve_items(Carver& c, const nh_model* m, int64_t n_items) {
  const int N = m->LX * m->LY;
  const int D = feature_dim(m);
  const int64_t images = n_items * ipi_of(m);
  const size_t tri = image_global_floats(m);
  ItemBuffers b;
  b.chunk = std::min<int64_t>(images, image_chunk(m));
  b.Y = c.take<float>((size_t)b.chunk * 2 * N * D);
  b.gtri = tri ? c.take<float>((size_t)b.chunk * tri) : nullptr;
  const size_t net = resconv_scratch_floats(m, b.chunk);
  b.net = net ? c.take<float>(net) : nullptr;
  b.img_sign = c.take<float>((size_t)images);
  b.img_logabs = c.take<double>((size_t)images);
  return b;
}

static int run_items(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                     const int32_t* item_hop, int64_t n_items, const ItemBuffers& b, float* sign,
                     double* logabs) {
  const int ipi = ipi_of(m);
  const int64_t images = n_items * ipi;
  for (int64_t i0 = 0; i0 < images; i0 += b.chunk) {
    const int64_t i1 = std::min(images, i0 + b.chunk);
    int rc;
    {
      ProfScope ps(st, NH_PROF_FEATURES);
      rc = launch_cnn(st, m, s, item_b, item_hop, (int)i0, (int)i1, ipi, b.Y, b.net);
    }
    if (rc != NH_OK) return rc;
    {
      ProfScope ps(st, NH_PROF_PAIRING_LINALG);
      rc = launch_image_amplitude(st, m, s, item_b, item_hop, (int)i0, (int)i1, ipi, b.Y, b.img_sign,
                                  b.img_logabs, b.gtri);
    }
    if (rc != NH_OK) return rc;
  }
  ProfScope ps(st, NH_PROF_OTHER);
  return launch_combine(st, item_hop, 0, (int)n_items, ipi, b.img_sign, b.img_logabs, sign, logabs);
}

static int z_of(double tprime) { return tprime != 0.0 ? 8 : 4; }

size_t image_amplitudes_bytes(const nh_model* m, int64_t batch) {
  Carver c(nullptr);
  carve_items(c, m, batch);
  return c.off;
}

int image_amplitudes(hipStream_t st, const nh_model* m, const int8_t* s, int64_t batch, void* ws, float** img_sign,
                     double** img_logabs, float* sign, double* logabs) {
  Carver c(ws);
  const ItemBuffers b = carve_items(c, m, batch);
  *img_sign = b.img_sign;
  *img_logabs = b.img_logabs;
  return run_items(st, m, s, nullptr, nullptr, batch, b, sign, logabs);
}

}  // namespace nh

using namespace nh;

extern "C" size_t nh_log_psi_workspace_size(const nh_model* model, int32_t batch) {
  if (check_model(model) != NH_OK || batch < 0) return 0;
  Carver c(nullptr);
  carve_items(c, model, batch);
  return c.off + 256;
}

extern "C" int nh_log_psi(void* stream, const nh_model* model, int32_t batch, const int8_t* s, float* sign,
                          double* logabs, void* workspace, size_t workspace_bytes) {
  int rc = check_model(model);
  if (rc != NH_OK) return rc;
  if (batch < 0) return NH_ERR_ARG;
  if (batch == 0) return NH_OK;
  if (!s || !sign || !logabs || !workspace) return NH_ERR_ARG;
  if ((int64_t)batch * ipi_of(model) > INT32_MAX / 2) return NH_ERR_SHAPE;
  if (workspace_bytes < nh_log_psi_workspace_size(model, batch)) return NH_ERR_WORKSPACE;
  Carver c((void*)align_up((size_t)workspace));
  ItemBuffers b = carve_items(c, model, batch);
  return run_items((hipStream_t)stream, model, s, nullptr, nullptr, batch, b, sign, logabs);
}

// ---- allocation-free projected amplitude: the workspace is a static device array of
// NH_NOALLOC_SCRATCH_BYTES (each GPU the library runs on holds its own copy until the process
// ends).  Host code never uses the array's name as an address: hipGetSymbolAddress gives each
// device's copy once, with that device current, and the kernels receive that pointer.
static __device__ __align__(256) char scratch[NH_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static int scratch_of_current_device(void** out) {
  int dev = -1;
  NH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return NH_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NH_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NH_OK;
}

int nh_log_psi_items(hipStream_t st, const nh_model* m, const int8_t* s, const int32_t* item_b,
                     const int32_t* item_hop, int n_items, float* sign, double* logabs, void* ws, size_t wsb) {
  if (wsb < nh_log_psi_workspace_size(m, n_items)) return NH_ERR_WORKSPACE;
  Carver c((void*)align_up((size_t)ws));
  ItemBuffers b = carve_items(c, m, n_items);
  return run_items(st, m, s, item_b, item_hop, n_items, b, sign, logabs);
}

extern "C" int nh_log_psi_noalloc(void* stream, const nh_model* model, int32_t batch, const int8_t* s, float* sign,
                                  double* logabs) {
  int rc = check_model(model);
  if (rc != NH_OK) return rc;
  if (batch < 0) return NH_ERR_ARG;
  if (batch == 0) return NH_OK;
  if (!s || !sign || !logabs) return NH_ERR_ARG;
  // the largest pass (configurations per nh_log_psi call) whose workspace fits the scratch;
  // workspace sizes grow with the pass, so a binary search finds it
  const size_t cap = NH_NOALLOC_SCRATCH_BYTES;
  if (nh_log_psi_workspace_size(model, 1) > cap) return NH_ERR_WORKSPACE;  // one configuration does not fit
  int32_t lo = 1, hi = batch;
  while (lo < hi) {
    const int32_t mid = lo + (hi - lo + 1) / 2;
    if (nh_log_psi_workspace_size(model, mid) <= cap) lo = mid;
    else hi = mid - 1;
  }
  void* ws = nullptr;
  rc = scratch_of_current_device(&ws);
  if (rc != NH_OK) return rc;
  const size_t two_n = 2 * (size_t)model->LX * model->LY;
  for (int32_t b0 = 0; b0 < batch; b0 += lo) {
    const int32_t nb = std::min(lo, batch - b0);
    rc = nh_log_psi(stream, model, nb, s + (size_t)b0 * two_n, sign + b0, logabs + b0, ws, cap);
    if (rc != NH_OK) return rc;
  }
  return NH_OK;
}

static size_t eloc_carve(Carver& c, const nh_model* m, double tprime, int32_t chunk, int32_t** count,
                         int32_t** item_b, int32_t** item_hop, float** coef, float** hsign, double** hlog,
                         ItemBuffers* ib) {
  const int max_hops = (m->n_up + m->n_dn) * z_of(tprime);
  const int64_t items = (int64_t)chunk * max_hops;
  *count = c.take<int32_t>(chunk);
  *item_b = c.take<int32_t>(items);
  *item_hop = c.take<int32_t>(items);
  *coef = c.take<float>(items);
  *hsign = c.take<float>(items);
  *hlog = c.take<double>(items);
  *ib = carve_items(c, m, items);
  return c.off;
}

extern "C" size_t nh_local_energy_workspace_size(const nh_model* model, double tprime, int32_t chunk) {
  if (check_model(model) != NH_OK || chunk <= 0) return 0;
  Carver c(nullptr);
  int32_t *a, *b, *h;
  float *co, *hs;
  double* hl;
  ItemBuffers ib;
  eloc_carve(c, model, tprime, chunk, &a, &b, &h, &co, &hs, &hl, &ib);
  return c.off + 256;
}

extern "C" int nh_local_energy(void* stream, const nh_model* model, double t, double tprime, double U,
                               int32_t batch, const int8_t* s, const float* sign_s, const double* logabs_s,
                               double* eloc, int32_t chunk, void* workspace, size_t workspace_bytes) {
  int rc = check_model(model);
  if (rc != NH_OK) return rc;
  if (batch < 0 || chunk <= 0) return NH_ERR_ARG;
  if (batch == 0) return NH_OK;
  if (!s || !sign_s || !logabs_s || !eloc || !workspace) return NH_ERR_ARG;
  const int max_hops = (model->n_up + model->n_dn) * z_of(tprime);
  if ((int64_t)chunk * max_hops * ipi_of(model) > INT32_MAX / 2) return NH_ERR_SHAPE;
  if (workspace_bytes < nh_local_energy_workspace_size(model, tprime, chunk)) return NH_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  Carver c((void*)align_up((size_t)workspace));
  int32_t *count, *item_b, *item_hop;
  float *coef, *hsign;
  double* hlog;
  ItemBuffers ib;
  eloc_carve(c, model, tprime, chunk, &count, &item_b, &item_hop, &coef, &hsign, &hlog, &ib);
  const int N = model->LX * model->LY;
  for (int b0 = 0; b0 < batch; b0 += chunk) {
    const int cs = std::min(chunk, batch - b0);
    {
      ProfScope ps(st, NH_PROF_OTHER);
      rc = launch_hops(st, model->LX, model->LY, (float)t, (float)tprime, z_of(tprime), cs, b0, s, max_hops,
                       count, item_b, item_hop, coef, nullptr, nullptr);
    }
    if (rc != NH_OK) return rc;
    rc = run_items(st, model, s, item_b, item_hop, (int64_t)cs * max_hops, ib, hsign, hlog);
    if (rc != NH_OK) return rc;
    ProfScope ps(st, NH_PROF_OTHER);
    rc = launch_eloc_reduce(st, N, U, cs, b0, s, sign_s, logabs_s, max_hops, count, coef, hsign, hlog, eloc);
    if (rc != NH_OK) return rc;
  }
  return NH_OK;
}

extern "C" int nh_hops(void* stream, int32_t LX, int32_t LY, double t, double tprime, int32_t batch,
                       int32_t n_occ, const int8_t* s, int32_t max_hops, int32_t* count, int32_t* out_from,
                       int32_t* out_to, float* out_coef) {
  if (LX < 3 || LY < 3 || batch < 0 || n_occ < 0) return NH_ERR_ARG;
  if (batch == 0) return NH_OK;
  if (!s || !count || !out_from || !out_to || !out_coef) return NH_ERR_ARG;
  if (max_hops < n_occ * z_of(tprime)) return NH_ERR_SHAPE;
  return launch_hops((hipStream_t)stream, LX, LY, (float)t, (float)tprime, z_of(tprime), batch, 0, s, max_hops,
                     count, nullptr, nullptr, out_coef, out_from, out_to);
}

extern "C" int nh_profile_enable(int32_t on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on != 0;
  return NH_OK;
}

extern "C" int nh_profile_query(int32_t kernel_class, int64_t* launches, double* total_ms) {
  if (!launches || !total_ms || kernel_class < 0 || kernel_class >= NH_PROF_CLASSES) return NH_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_prof_mu);
  int64_t n = 0;
  double tot = 0.0;
  for (auto& r : g_prof) {
    if (r.kid != kernel_class) continue;
    NH_HIP(hipEventSynchronize(r.b));
    float ms = 0.f;
    NH_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    tot += ms;
    ++n;
  }
  *launches = n;
  *total_ms = tot;
  return NH_OK;
}

extern "C" int nh_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  int rc = NH_OK;
  for (auto& r : g_prof) {
    if (hipEventSynchronize(r.b) != hipSuccess) rc = NH_ERR_HIP;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_prof.clear();
  return rc;
}
