// C ABI of the dwmi library (declared in include/dwmi.h): parameter ingestion, device weight layouts,
// workspace management and the three drivers (log|psi|, local energy, MCMC).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dwmi.h"
#include "kernels.h"

using namespace dwmi;

namespace {

thread_local std::string g_err;

dw_status fail(dw_status code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                                    \
  do {                                                                                                   \
    hipError_t _e = (expr);                                                                              \
    if (_e != hipSuccess) return fail(DW_ERR_HIP, std::string(#expr) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      bytes = 0;
    }
    hipError_t e = hipMalloc(&p, need);
    if (e == hipSuccess) bytes = need;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  float* f() const { return (float*)p; }
  uint16_t* h() const { return (uint16_t*)p; }
  int32_t* i() const { return (int32_t*)p; }
  uint32_t* u() const { return (uint32_t*)p; }
};

struct Leaf {
  std::vector<float> data;
  std::vector<int64_t> shape;
  bool set = false;
};

uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf2f(uint16_t b) {
  uint32_t u = ((uint32_t)b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

std::string normalise(const char* path) {
  std::string p(path ? path : "");
  std::string out;
  size_t start = 0;
  while (start <= p.size()) {
    size_t e = p.find('/', start);
    if (e == std::string::npos) e = p.size();
    std::string seg = p.substr(start, e - start);
    if (!seg.empty() && seg != "~") {
      if (!out.empty()) out += "/";
      out += seg;
    }
    start = e + 1;
  }
  return out;
}

bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0 &&
         (s.size() == suf.size() || s[s.size() - suf.size() - 1] == '/');
}

}  // namespace

struct dw_model {
  dw_model_desc d{};
  int n = 0, u = 0, dn = 0, I = 0, D = 0, W1 = 0, W2 = 0, L = 0, NO = 0, NOp = 0;
  std::vector<int> Kh, Ke, Kx, Km;
  std::map<std::string, Leaf> leaves;  // key "<module>:<leaf>"
  bool dirty = true;
  int chunk = 0;
  GemmPrec prec = GEMM_F32;
  // device weights
  std::vector<DevBuf> Wmain, Wmu, bias, Wmain_hi, Wmain_lo, Wmu_hi, Wmu_lo;
  DevBuf W2e, b2e, Wo[2], Wo_hi[2], Wo_lo[2], kappa, wenv;
  // workspace
  std::vector<DevBuf> X;
  DevBuf MU, V, H, O, A, ldd, sdd, gd, lapd, dfl, rprop, lpprop, cnt;
  std::vector<std::pair<std::string, std::vector<int64_t>>> expected;
  // measurement hooks
  bool timing = false;
  struct TimerRec {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<TimerRec> timers;
  std::vector<hipEvent_t> event_pool;

  ~dw_model() {
    for (auto& t : timers) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    for (auto e : event_pool) (void)hipEventDestroy(e);
    for (auto* v : {&Wmain, &Wmu, &bias, &Wmain_hi, &Wmain_lo, &Wmu_hi, &Wmu_lo, &X})
      for (auto& b : *v) b.release();
    for (DevBuf* b : {&W2e, &b2e, &Wo[0], &Wo[1], &Wo_hi[0], &Wo_hi[1], &Wo_lo[0], &Wo_lo[1], &kappa, &wenv, &MU, &V,
                      &H, &O, &A, &ldd, &sdd, &gd, &lapd, &dfl, &rprop, &lpprop, &cnt})
      b->release();
  }
};

namespace {

int pad_to(int x, int mult) { return (x + mult - 1) / mult * mult; }

std::string leaf_key(const std::string& module, const std::string& leaf) { return module + ":" + leaf; }

// Identify a leaf from its (normalised) module path.  Returns "" if unknown.
std::string classify(const dw_model* m, const std::string& path, const std::string& leaf) {
  for (int l = 0; l < m->L; ++l) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "h_el_%d/linear_0", l);
    if (ends_with(path, buf) && (leaf == "w" || leaf == "b")) return leaf_key(buf, leaf);
    if (l < m->L - 1) {
      std::snprintf(buf, sizeof buf, "h_el_el_%d/linear_0", l);
      if (ends_with(path, buf) && (leaf == "w" || leaf == "b")) return leaf_key(buf, leaf);
    }
  }
  if (ends_with(path, "bf_up/linear_0") && leaf == "w") return leaf_key("bf_up/linear_0", "w");
  if (ends_with(path, "bf_dn/linear_0") && leaf == "w") return leaf_key("bf_dn/linear_0", "w");
  if (ends_with(path, "envelope_orbitals") &&
      (leaf == "alpha_up" || leaf == "alpha_dn" || leaf == "weights_up" || leaf == "weights_dn"))
    return leaf_key("envelope_orbitals", leaf);
  return "";
}

template <typename T>
dw_status upload(DevBuf& b, const std::vector<T>& host) {
  HIP_TRY(b.ensure(host.size() * sizeof(T)));
  HIP_TRY(hipMemcpy(b.p, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice));
  return DW_OK;
}

dw_status upload_split(DevBuf& hi, DevBuf& lo, const std::vector<float>& w) {
  std::vector<uint16_t> h(w.size()), l(w.size());
  for (size_t k = 0; k < w.size(); ++k) {
    h[k] = f2bf(w[k]);
    l[k] = f2bf(w[k] - bf2f(h[k]));
  }
  dw_status st = upload(hi, h);
  if (st) return st;
  return upload(lo, l);
}

// Build the device layouts of all weights from the host leaves.
dw_status prepare(dw_model* m) {
  if (!m->dirty) return DW_OK;
  for (auto& e : m->expected)
    if (!m->leaves[e.first].set) return fail(DW_ERR_STATE, "parameter not set: " + e.first);
  const int W1 = m->W1, W2 = m->W2;
  const bool split = m->prec != GEMM_F32;
  for (int l = 0; l < m->L; ++l) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "h_el_%d/linear_0", l);
    const Leaf& w = m->leaves[leaf_key(buf, "w")];
    const Leaf& b = m->leaves[leaf_key(buf, "b")];
    const int Kh = m->Kh[l], Ke = m->Ke[l], Kx = m->Kx[l], Km = m->Km[l];
    std::vector<float> wm((size_t)W1 * Kx, 0.f), wu((size_t)W1 * Km, 0.f);
    for (int o = 0; o < W1; ++o) {
      for (int k = 0; k < Kh; ++k) wm[(size_t)o * Kx + k] = w.data[(size_t)k * W1 + o];
      for (int k = 0; k < 2 * Ke; ++k) wm[(size_t)o * Kx + Kh + k] = w.data[(size_t)(3 * Kh + k) * W1 + o];
      for (int k = 0; k < 2 * Kh; ++k) wu[(size_t)o * Km + k] = w.data[(size_t)(Kh + k) * W1 + o];
    }
    dw_status st;
    if ((st = upload(m->Wmain[l], wm))) return st;
    if ((st = upload(m->Wmu[l], wu))) return st;
    if ((st = upload(m->bias[l], b.data))) return st;
    if (split) {
      if ((st = upload_split(m->Wmain_hi[l], m->Wmain_lo[l], wm))) return st;
      if ((st = upload_split(m->Wmu_hi[l], m->Wmu_lo[l], wu))) return st;
    }
  }
  {
    const int NL2 = m->L - 1 > 0 ? m->L - 1 : 1;
    std::vector<float> w2((size_t)NL2 * 32 * 32, 0.f), b2((size_t)NL2 * 32, 0.f);
    for (int l = 0; l < m->L - 1; ++l) {
      char buf[64];
      std::snprintf(buf, sizeof buf, "h_el_el_%d/linear_0", l);
      const Leaf& w = m->leaves[leaf_key(buf, "w")];
      const Leaf& b = m->leaves[leaf_key(buf, "b")];
      const int kin = (l == 0) ? 4 : W2;
      for (int k = 0; k < kin; ++k)
        for (int f = 0; f < W2; ++f) w2[((size_t)l * 32 + k) * 32 + f] = w.data[(size_t)k * W2 + f];
      for (int f = 0; f < W2; ++f) b2[(size_t)l * 32 + f] = b.data[f];
    }
    dw_status st;
    if ((st = upload(m->W2e, w2))) return st;
    if ((st = upload(m->b2e, b2))) return st;
  }
  const char* bfk[2] = {"bf_up/linear_0", "bf_dn/linear_0"};
  const char* ak[2] = {"alpha_up", "alpha_dn"};
  const char* wk[2] = {"weights_up", "weights_dn"};
  std::vector<float> kap((size_t)2 * m->I * m->NO), wen((size_t)2 * m->I * m->NO);
  for (int s = 0; s < 2; ++s) {
    const Leaf& w = m->leaves[leaf_key(bfk[s], "w")];
    std::vector<float> wo((size_t)m->NOp * W1, 0.f);
    for (int c = 0; c < m->NO; ++c)
      for (int k = 0; k < W1; ++k) wo[(size_t)c * W1 + k] = w.data[(size_t)k * m->NO + c];
    dw_status st;
    if ((st = upload(m->Wo[s], wo))) return st;
    if (split && (st = upload_split(m->Wo_hi[s], m->Wo_lo[s], wo))) return st;
    const Leaf& a = m->leaves[leaf_key("envelope_orbitals", ak[s])];
    const Leaf& we = m->leaves[leaf_key("envelope_orbitals", wk[s])];
    for (size_t q = 0; q < (size_t)m->I * m->NO; ++q) {
      const double x = a.data[q];
      const double sp = x > 30.0 ? x : std::log1p(std::exp(x));  // softplus
      kap[(size_t)s * m->I * m->NO + q] = (float)sp;
      wen[(size_t)s * m->I * m->NO + q] = we.data[q];
    }
  }
  dw_status st;
  if ((st = upload(m->kappa, kap))) return st;
  if ((st = upload(m->wenv, wen))) return st;
  HIP_TRY(hipDeviceSynchronize());
  m->dirty = false;
  return DW_OK;
}

int rows_per_electron(int C) { return C > 1 ? pad_to(C, 16) : 1; }

// Size the workspace for B walkers with C channels.
dw_status ensure_ws(dw_model* m, long B, int C) {
  const int Cp = rows_per_electron(C);
  const long rows = B * m->n * Cp;
  for (int l = 0; l < m->L; ++l) HIP_TRY(m->X[l].ensure((size_t)rows * m->Kx[l] * 4));
  int kmmax = 0;
  for (int l = 0; l < m->L; ++l) kmmax = std::max(kmmax, m->Km[l]);
  HIP_TRY(m->MU.ensure((size_t)B * Cp * kmmax * 4));
  HIP_TRY(m->V.ensure((size_t)B * Cp * m->W1 * 4));
  HIP_TRY(m->H.ensure((size_t)rows * m->W1 * 4));
  HIP_TRY(m->O.ensure((size_t)rows * m->NO * 4));
  HIP_TRY(m->A.ensure((size_t)B * m->D * C * m->n * m->n * 4));
  HIP_TRY(m->ldd.ensure((size_t)B * m->D * 4));
  HIP_TRY(m->sdd.ensure((size_t)B * m->D * 4));
  HIP_TRY(m->dfl.ensure((size_t)B * m->D * 4));
  if (C > 1) {
    HIP_TRY(m->gd.ensure((size_t)B * m->D * 3 * m->n * 4));
    HIP_TRY(m->lapd.ensure((size_t)B * m->D * 4));
  } else {
    HIP_TRY(m->gd.ensure(4));
    HIP_TRY(m->lapd.ensure(4));
  }
  return DW_OK;
}

size_t ws_bytes(const dw_model* m, long B, int C) {
  const int Cp = rows_per_electron(C);
  const long rows = B * m->n * Cp;
  size_t tot = 0;
  int kmmax = 0;
  for (int l = 0; l < m->L; ++l) {
    tot += (size_t)rows * m->Kx[l] * 4;
    kmmax = std::max(kmmax, m->Km[l]);
  }
  tot += (size_t)B * Cp * kmmax * 4 + (size_t)B * Cp * m->W1 * 4;
  tot += (size_t)rows * m->W1 * 4 + (size_t)rows * m->NO * 4;
  tot += (size_t)B * m->D * C * m->n * m->n * 4;
  tot += (size_t)B * m->D * 4 * 3;
  if (C > 1) tot += (size_t)B * m->D * (3 * m->n + 1) * 4;
  return tot;
}

hipEvent_t get_event(dw_model* m) {
  if (!m->event_pool.empty()) {
    hipEvent_t e = m->event_pool.back();
    m->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Bracket a region with events on stream s when timing is on.
struct Timed {
  dw_model* m;
  int kind;
  hipStream_t s;
  hipEvent_t a = nullptr;
  // kind < 0: not timed
  Timed(dw_model* m_, int kind_, hipStream_t s_) : m(m_), kind(kind_), s(s_) {
    if (m->timing && kind >= 0) {
      a = get_event(m);
      (void)hipEventRecord(a, s);
    }
  }
  ~Timed() {
    if (a) {
      hipEvent_t b = get_event(m);
      (void)hipEventRecord(b, s);
      m->timers.push_back({a, b, kind});
    }
  }
};

// One pass of the wavefunction over B walkers with C channels (C == 1: value; C == 3n+2: Laplacian).
dw_status run_pass(dw_model* m, const float* r, long B, const float* R, const float* Z, int C, hipStream_t s,
                   float* out_lp, float* out_phase, float* out_eloc, uint32_t* out_flags) {
  dw_status st = ensure_ws(m, B, C);
  if (st) return st;
  const int Cp = rows_per_electron(C);
  Shape sh{(int)B, m->n, m->u, m->I, m->D, C, Cp, m->W1, m->W2, m->L};
  const long rows = B * m->n * Cp;
  Timed pass_timer(m, C > 1 ? DW_TIMER_LOCAL_ENERGY : DW_TIMER_FORWARD, s);
  const GemmPrec pr = m->prec;
  std::vector<float*> X(m->L);
  for (int l = 0; l < m->L; ++l) X[l] = m->X[l].f();
  launch_features0(r, R, X[0], m->Kx[0], sh, s);
  launch_twoel(r, X.data(), m->Kx.data(), m->Kh.data(), m->W2e.f(), m->b2e.f(), sh, s);
  for (int l = 0; l < m->L; ++l) {
    launch_spin_mean(X[l], m->Kx[l], m->Kh[l], m->MU.f(), m->Km[l], sh, s);
    GemmEpi ev;
    ev.bias = m->bias[l].f();
    ev.C = Cp;
    ev.nC = m->n * Cp;
    gemm(pr, m->MU.f(), m->Km[l], m->Wmu[l].f(), m->Wmu_hi[l].h(), m->Wmu_lo[l].h(), m->Km[l], m->V.f(), m->W1,
         (int)(B * Cp), m->W1, m->Km[l], ev, s);
    const bool last = (l == m->L - 1);
    float* outp = last ? m->H.f() : X[l + 1];
    const int ldo = last ? m->W1 : m->Kx[l + 1];
    if (C == 1) {
      GemmEpi em;
      em.V = m->V.f();
      em.ldv = m->W1;
      em.C = 1;
      em.nC = m->n;
      em.act_tanh = 1;
      em.permute = last ? 1 : 0;
      em.n = m->n;
      em.u = m->u;
      em.B = (int)B;
      gemm(pr, X[l], m->Kx[l], m->Wmain[l].f(), m->Wmain_hi[l].h(), m->Wmain_lo[l].h(), m->Kx[l], outp, ldo,
           (int)rows, m->W1, m->Kx[l], em, s);
    } else {
      Timed t(m, l >= 1 ? DW_TIMER_GEMM_1EL_LAP : -1, s);
      gemm_lap(pr, X[l], m->Kx[l], m->Wmain[l].f(), m->Wmain_hi[l].h(), m->Wmain_lo[l].h(), m->Kx[l], m->V.f(),
               m->W1, outp, ldo, last ? 1 : 0, sh, s);
    }
  }
  // orbital GEMMs on the spin-major rows
  const long up_rows = B * m->u * Cp, dn_rows = B * m->dn * Cp;
  GemmEpi e0;
  gemm(pr, m->H.f(), m->W1, m->Wo[0].f(), m->Wo_hi[0].h(), m->Wo_lo[0].h(), m->W1, m->O.f(), m->NO, (int)up_rows,
       m->NO, m->W1, e0, s);
  gemm(pr, m->H.f() + up_rows * m->W1, m->W1, m->Wo[1].f(), m->Wo_hi[1].h(), m->Wo_lo[1].h(), m->W1,
       m->O.f() + up_rows * m->NO, m->NO, (int)dn_rows, m->NO, m->W1, e0, s);
  launch_envelope(m->O.f(), m->NO, r, R, m->kappa.f(), m->wenv.f(), m->A.f(), sh, s);
  {
    Timed t(m, C > 1 ? DW_TIMER_DET_LAP : -1, s);
    launch_det(m->A.f(), m->ldd.f(), m->sdd.f(), m->gd.f(), m->lapd.f(), m->dfl.u(), m->d.fp64_cond_threshold, sh, s);
  }
  launch_combine(m->ldd.f(), m->sdd.f(), m->gd.f(), m->lapd.f(), m->dfl.u(), r, R, Z, out_lp, out_phase, out_eloc,
                 out_flags, sh, s);
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

}  // namespace

extern "C" {

const char* dw_last_error(void) { return g_err.c_str(); }
const char* dw_version(void) { return "dwmi 0.1 (gfx950)"; }

dw_status dw_model_set_timing(dw_model* m, int32_t enable) {
  if (!m) return fail(DW_ERR_INVALID, "null model");
  for (auto& t : m->timers) {
    m->event_pool.push_back(t.a);
    m->event_pool.push_back(t.b);
  }
  m->timers.clear();
  m->timing = enable != 0;
  return DW_OK;
}

dw_status dw_model_timer_query(dw_model* m, int32_t kind, double* total_ms, int32_t* launches) {
  if (!m || !total_ms || !launches) return fail(DW_ERR_INVALID, "null argument");
  double tot = 0.0;
  int cnt = 0;
  for (auto& t : m->timers) {
    if (t.kind != kind) continue;
    HIP_TRY(hipEventSynchronize(t.b));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
    tot += ms;
    ++cnt;
  }
  *total_ms = tot;
  *launches = cnt;
  return DW_OK;
}

dw_status dw_model_create(const dw_model_desc* desc, dw_model** out) {
  if (!desc || !out) return fail(DW_ERR_INVALID, "null argument");
  const dw_model_desc& d = *desc;
  if (d.n_el < 2 || d.n_el > 64) return fail(DW_ERR_INVALID, "n_el must be in [2, 64]");
  if (d.n_up < 1 || d.n_up >= d.n_el) return fail(DW_ERR_INVALID, "need 1 <= n_up < n_el (both spin blocks non-empty)");
  if (d.n_ion < 1 || d.n_ion > 64) return fail(DW_ERR_INVALID, "n_ion must be in [1, 64]");
  if (d.n_det < 1) return fail(DW_ERR_INVALID, "n_det must be >= 1");
  if (d.width2 != 32) return fail(DW_ERR_INVALID, "the fused 2-el kernel requires width2 == 32");
  if (d.width1 < 32 || d.width1 % 128 != 0) return fail(DW_ERR_INVALID, "width1 must be a multiple of 128");
  if (d.n_iter < 2 || d.n_iter > 4) return fail(DW_ERR_INVALID, "n_iter must be in [2, 4]");
  if (d.precision < DW_FP32 || d.precision > DW_BF16) return fail(DW_ERR_INVALID, "unknown precision");
  dw_model* m = new dw_model();
  m->d = d;
  m->n = d.n_el;
  m->u = d.n_up;
  m->dn = d.n_el - d.n_up;
  m->I = d.n_ion;
  m->D = d.n_det;
  m->W1 = d.width1;
  m->W2 = d.width2;
  m->L = d.n_iter;
  m->NO = m->D * m->n;
  m->NOp = pad_to(m->NO, 128);
  m->prec = (GemmPrec)d.precision;
  for (int l = 0; l < m->L; ++l) {
    const int Kh = (l == 0) ? 4 * m->I : m->W1;
    const int Ke = (l == 0) ? 4 : m->W2;
    m->Kh.push_back(Kh);
    m->Ke.push_back(Ke);
    m->Kx.push_back(pad_to(Kh + 2 * Ke, 32));
    m->Km.push_back(pad_to(2 * Kh, 32));
  }
  m->Wmain.resize(m->L);
  m->Wmu.resize(m->L);
  m->bias.resize(m->L);
  m->Wmain_hi.resize(m->L);
  m->Wmain_lo.resize(m->L);
  m->Wmu_hi.resize(m->L);
  m->Wmu_lo.resize(m->L);
  m->X.resize(m->L);
  for (int l = 0; l < m->L; ++l) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "h_el_%d/linear_0", l);
    m->expected.push_back({leaf_key(buf, "w"), {3 * m->Kh[l] + 2 * m->Ke[l], m->W1}});
    m->expected.push_back({leaf_key(buf, "b"), {m->W1}});
    if (l < m->L - 1) {
      std::snprintf(buf, sizeof buf, "h_el_el_%d/linear_0", l);
      m->expected.push_back({leaf_key(buf, "w"), {l == 0 ? 4 : m->W2, m->W2}});
      m->expected.push_back({leaf_key(buf, "b"), {m->W2}});
    }
  }
  m->expected.push_back({leaf_key("bf_up/linear_0", "w"), {m->W1, m->NO}});
  m->expected.push_back({leaf_key("bf_dn/linear_0", "w"), {m->W1, m->NO}});
  for (const char* k : {"alpha_up", "alpha_dn", "weights_up", "weights_dn"})
    m->expected.push_back({leaf_key("envelope_orbitals", k), {m->I, m->NO}});
  *out = m;
  return DW_OK;
}

dw_status dw_model_destroy(dw_model* m) {
  delete m;
  return DW_OK;
}

dw_status dw_model_set_param(dw_model* m, const char* module_path, const char* leaf, const float* host_data,
                             const int64_t* shape, int32_t ndim) {
  if (!m || !module_path || !leaf || !host_data || (!shape && ndim > 0)) return fail(DW_ERR_INVALID, "null argument");
  const std::string key = classify(m, normalise(module_path), leaf);
  if (key.empty()) return fail(DW_ERR_INVALID, std::string("unknown parameter ") + module_path + "/" + leaf);
  const std::vector<int64_t>* exp_shape = nullptr;
  for (auto& e : m->expected)
    if (e.first == key) exp_shape = &e.second;
  std::vector<int64_t> sh(shape, shape + ndim);
  if (!exp_shape || sh != *exp_shape) {
    std::string msg = "shape mismatch for " + key + ": got [";
    for (auto v : sh) msg += std::to_string(v) + ",";
    msg += "] expected [";
    if (exp_shape)
      for (auto v : *exp_shape) msg += std::to_string(v) + ",";
    return fail(DW_ERR_INVALID, msg + "]");
  }
  size_t count = 1;
  for (auto v : sh) count *= (size_t)v;
  Leaf& lf = m->leaves[key];
  lf.data.assign(host_data, host_data + count);
  lf.shape = sh;
  lf.set = true;
  m->dirty = true;
  return DW_OK;
}

int32_t dw_model_ready(const dw_model* m) {
  if (!m) return 0;
  for (auto& e : m->expected) {
    auto it = m->leaves.find(e.first);
    if (it == m->leaves.end() || !it->second.set) return 0;
  }
  return 1;
}

dw_status dw_model_set_chunk(dw_model* m, int32_t max_walkers_per_pass) {
  if (!m || max_walkers_per_pass < 0) return fail(DW_ERR_INVALID, "bad chunk");
  m->chunk = max_walkers_per_pass;
  return DW_OK;
}

size_t dw_workspace_size(const dw_model* m, int32_t n_walkers, int32_t mode) {
  if (!m || n_walkers <= 0) return 0;
  const int C = mode ? 3 * m->n + 2 : 1;
  long B = n_walkers;
  if (mode && m->chunk > 0 && B > m->chunk) B = m->chunk;
  return ws_bytes(m, B, C);
}

dw_status dw_logpsi(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                    float* out_log_psi_sqr, float* out_phase, dw_stream stream) {
  if (!m || !r || !R || !Z || !out_log_psi_sqr) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  dw_status st = prepare(m);
  if (st) return st;
  return run_pass(m, r, n_walkers, R, Z, 1, (hipStream_t)stream, out_log_psi_sqr, out_phase, nullptr, nullptr);
}

dw_status dw_local_energy(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                          float* out_eloc, float* out_log_psi_sqr, uint32_t* out_flags, dw_stream stream) {
  if (!m || !r || !R || !Z || !out_eloc) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  dw_status st = prepare(m);
  if (st) return st;
  const int C = 3 * m->n + 2;
  const long chunk = m->chunk > 0 ? m->chunk : n_walkers;
  for (long b0 = 0; b0 < n_walkers; b0 += chunk) {
    const long nb = std::min<long>(chunk, n_walkers - b0);
    st = run_pass(m, r + b0 * m->n * 3, nb, R, Z, C, (hipStream_t)stream,
                  out_log_psi_sqr ? out_log_psi_sqr + b0 : nullptr, nullptr, out_eloc + b0,
                  out_flags ? out_flags + b0 : nullptr);
    if (st) return st;
  }
  return DW_OK;
}

dw_status dw_mcmc_steps(dw_model* m, float* r, int32_t n_walkers, const float* R, const float* Z, float* log_psi_sqr,
                        int32_t* walker_age, const float* normals, const float* uniforms, int32_t n_steps,
                        dw_chain_scalars* chain, int32_t max_age, int32_t stepsize_update_interval,
                        float target_acceptance, float min_stepsize, float max_stepsize, float* out_acc,
                        uint32_t* out_flags, dw_stream stream) {
  if (!m || !r || !R || !Z || !log_psi_sqr || !walker_age || !chain) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0 || n_steps < 0) return fail(DW_ERR_INVALID, "bad sizes");
  if (n_steps > 0 && (!normals || !uniforms)) return fail(DW_ERR_INVALID, "draws required");
  dw_status st = prepare(m);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  const long B = n_walkers;
  const int cnt3 = (int)(B * m->n * 3);
  HIP_TRY(m->rprop.ensure((size_t)cnt3 * 4));
  HIP_TRY(m->lpprop.ensure((size_t)B * 4));
  HIP_TRY(m->cnt.ensure((size_t)std::max(n_steps, 1) * 4));
  // log_psi_sqr is re-evaluated on entry (mcmc.py:396)
  st = run_pass(m, r, B, R, Z, 1, s, log_psi_sqr, nullptr, nullptr, nullptr);
  if (st) return st;
  if (n_steps == 0) return DW_OK;
  launch_fill_i32(m->cnt.i(), 0, n_steps, s);
  for (int k = 0; k < n_steps; ++k) {
    launch_propose(r, normals + (size_t)k * cnt3, chain, m->rprop.f(), cnt3, s);
    st = run_pass(m, m->rprop.f(), B, R, Z, 1, s, m->lpprop.f(), nullptr, nullptr, nullptr);
    if (st) return st;
    launch_accept(r, m->rprop.f(), log_psi_sqr, m->lpprop.f(), uniforms + (size_t)k * B, walker_age, max_age,
                  m->cnt.i() + k, out_flags, (int)B, m->n, s);
    launch_chain_update(chain, m->cnt.i() + k, (int)B, out_acc ? out_acc + k : nullptr, stepsize_update_interval,
                        target_acceptance, min_stepsize, max_stepsize, s);
  }
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

}  // extern "C"
