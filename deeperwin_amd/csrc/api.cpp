// C ABI of the dwmi library (declared in include/dwmi.h): parameter ingestion, device weight layouts,
// workspace management and the three drivers (log|psi|, local energy, MCMC).  The wavefunction pass is a
// template over the activation type: float (FP32 / BF16 GEMM forms) or double (FP64 form and the FP64
// recompute of badly conditioned walkers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "api_internal.h"
#include "dwmi.h"
#include "grad.h"
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
  template <typename T>
  T* as() const {
    return (T*)p;
  }
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

int pad_to(int x, int mult) { return (x + mult - 1) / mult * mult; }
int rows_per_electron(int C) { return C > 1 ? pad_to(C, 16) : 1; }

// Device weight layouts for one activation type.
struct Weights {
  bool ready = false;
  std::vector<DevBuf> Wmain, Wmu, bias, Wmain_hi, Wmain_lo, Wmu_hi, Wmu_lo;
  std::vector<DevBuf> WmainT, WmuT;  // transposed copies for the gradient pass ([pad128(K)][W1])
  DevBuf W2e, b2e, Wo[2], Wo_hi[2], Wo_lo[2], WoT[2], kappa, wenv;
  // dpe4: pair-stream sets, el-ion stream, ion maps, ion table; h_map GEMM operands per iteration
  DevBuf Wp, bp, Wei, bei, Whim, bhim, table;
  std::vector<DevBuf> WhmT, WhmT_hi, WhmT_lo, bhm, WhmB;  // WhmB: h_map w [pad128(Kh)][32] (gradient pass)
  void release() {
    for (auto* v : {&Wmain, &Wmu, &bias, &Wmain_hi, &Wmain_lo, &Wmu_hi, &Wmu_lo, &WmainT, &WmuT, &WhmT, &WhmT_hi,
                    &WhmT_lo, &bhm, &WhmB})
      for (auto& b : *v) b.release();
    for (DevBuf* b : {&W2e, &b2e, &Wo[0], &Wo[1], &Wo_hi[0], &Wo_hi[1], &Wo_lo[0], &Wo_lo[1], &WoT[0], &WoT[1], &kappa,
                      &wenv, &Wp, &bp, &Wei, &bei, &Whim, &bhim, &table})
      b->release();
    ready = false;
  }
};

// Device scratch of one wavefunction pass over B walkers with C channels (activation type T): the
// activation buffers of every layer and the per-determinant results.  All of it is carved, 256-byte
// aligned, out of one region: the model's own arena, a caller-owned workspace, or a fixed scratch array.
template <typename T>
struct PassBufs {
  T* X[4] = {nullptr, nullptr, nullptr, nullptr};
  T *MU = nullptr, *V = nullptr, *Z = nullptr, *H = nullptr, *O = nullptr, *A = nullptr;
  T *WM = nullptr, *HM = nullptr, *HIM = nullptr;  // dpe4: mapped pairs, h_map rows, ion maps
  double *ld = nullptr, *sd = nullptr, *g = nullptr, *lap = nullptr;
  float* dcond = nullptr;
};

// Bump allocator over a region; with base == nullptr it only counts bytes.
struct Carver {
  char* base = nullptr;
  size_t off = 0;
  template <typename U>
  U* take(size_t count) {
    U* q = base ? (U*)(base + off) : nullptr;
    off += (count * sizeof(U) + 255) & ~(size_t)255;
    return q;
  }
};

}  // namespace

struct dw_model {
  dw_model_desc d{};
  int n = 0, u = 0, dn = 0, I = 0, D = 0, W1 = 0, W2 = 0, L = 0, NO = 0, NOp = 0, NOk = 0;
  std::vector<int> Kh, Ke, Kx, Km;
  std::vector<int> Kt0, Kt;  // columns [Kt0, Kt0 + Kt) of X_l: the 2-el block (pair means, or e_ee | e_ei)
  bool dpe4 = false;         // DW_EMBEDDING_DPE4
  int z_min = 1, types = 0;  // dpe4 ion table
  std::map<std::string, Leaf> leaves;  // key "<module>:<leaf>"
  std::vector<std::pair<std::string, std::vector<int64_t>>> expected;
  bool dirty = true;
  int chunk = 0;
  dw_tuning tune{-1, -1, 0, -1, -1, 0, 0};
  void* ext_ws = nullptr;  // caller-owned workspace (dw_model_set_workspace)
  size_t ext_ws_bytes = 0;
  GemmPrec prec = GEMM_F32;
  Weights wf, wd;    // float and double layouts
  DevBuf arena_f, arena_d;  // pass scratch (PassBufs) of the float and double activation types
  DevBuf grad_arena;        // scratch of the parameter-gradient pass
  DevBuf kappa_w;           // per-walker conditioning of the last pass (dw_model_conditioning)
  long kappa_n = 0;         // walkers covered by kappa_w
  DevBuf rprop, lpprop, logq, cnt;
  int prop_kind = DW_PROPOSAL_NORMAL;  // dw_model_set_proposal
  float prop_rmin = 0.1f, prop_rmax = 1.0f, prop_scale = 1.0f;
  DevBuf sub_r, sub_e, sub_lp, sub_idx, own_flags;  // FP64 recompute of flagged walkers
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
    wf.release();
    wd.release();
    for (DevBuf* b : {&arena_f, &arena_d, &grad_arena, &kappa_w, &rprop, &lpprop, &logq, &cnt, &sub_r, &sub_e,
                      &sub_lp, &sub_idx, &own_flags})
      b->release();
  }
};

namespace {

std::string leaf_key(const std::string& module, const std::string& leaf) { return module + ":" + leaf; }

// Identify a leaf from its (normalised) module path.  Returns "" if unknown.
std::string classify(const dw_model* m, const std::string& path, const std::string& leaf) {
  const bool wb = leaf == "w" || leaf == "b";
  for (int l = 0; l < m->L; ++l) {
    char buf[96];
    std::snprintf(buf, sizeof buf, "h_el_%d/linear_0", l);
    if (ends_with(path, buf) && wb) return leaf_key(buf, leaf);
    if (m->dpe4) {
      for (const char* mod : {"w_same", "w_diff", "h_map", "h_ion_map"}) {
        std::snprintf(buf, sizeof buf, "symm_features_%d/convolutional_features/%s/linear_0", l, mod);
        if (ends_with(path, buf) && wb) return leaf_key(buf, leaf);
      }
      if (l < m->L - 1)
        for (const char* mod : {"h_same", "h_diff", "h_el_ion"}) {
          std::snprintf(buf, sizeof buf, "%s_%d/linear_0", mod, l);
          if (ends_with(path, buf) && wb) return leaf_key(buf, leaf);
        }
    } else if (l < m->L - 1) {
      std::snprintf(buf, sizeof buf, "h_el_el_%d/linear_0", l);
      if (ends_with(path, buf) && wb) return leaf_key(buf, leaf);
    }
  }
  if (m->dpe4 && ends_with(path, "input/h_ion") && leaf == "embeddings") return leaf_key("input/h_ion", leaf);
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

template <typename T>
dw_status upload_as(DevBuf& b, const std::vector<float>& host) {
  std::vector<T> v(host.begin(), host.end());
  return upload(b, v);
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

// dpe4 weights: the pair-stream sets Wp [2L-1][2][32][32] (maps w_same/w_diff of every iteration, then the
// stream layers h_same/h_diff), the el-ion stream Wei [L-1][32][32], the ion maps Whim [L][32][32] (output
// columns beyond the iteration's width zero), the ion table, and per iteration the h_map GEMM operand
// WhmT[l] [128][pad32(Kh)] (row = output feature) with its bias.  Fan-in rows beyond a layer's width are 0.
dw_status prepare_dpe4(dw_model* m, Weights& W, bool dbl, bool split) {
  const int L = m->L, E = 32, NS = 2 * L - 1;
  auto up = [&](DevBuf& b, const std::vector<float>& v) { return dbl ? upload_as<double>(b, v) : upload(b, v); };
  auto leafp = [&](const std::string& mod, const char* leaf) -> const Leaf& { return m->leaves[leaf_key(mod, leaf)]; };
  std::vector<float> wp((size_t)NS * 2 * E * E, 0.f), bpv((size_t)NS * 2 * E, 0.f);
  for (int q = 0; q < NS; ++q) {
    const int l = q < L ? q : q - L;
    const int kin = l == 0 ? 1 : E;
    for (int sp = 0; sp < 2; ++sp) {
      char buf[96];
      if (q < L)
        std::snprintf(buf, sizeof buf, "symm_features_%d/convolutional_features/%s/linear_0", l,
                      sp == 0 ? "w_same" : "w_diff");
      else
        std::snprintf(buf, sizeof buf, "%s_%d/linear_0", sp == 0 ? "h_same" : "h_diff", l);
      const Leaf& w = leafp(buf, "w");
      const Leaf& b = leafp(buf, "b");
      for (int k = 0; k < kin; ++k)
        for (int f = 0; f < E; ++f) wp[(((size_t)q * 2 + sp) * E + k) * E + f] = w.data[(size_t)k * E + f];
      for (int f = 0; f < E; ++f) bpv[((size_t)q * 2 + sp) * E + f] = b.data[f];
    }
  }
  const int NE = L - 1 > 0 ? L - 1 : 1;
  std::vector<float> wei((size_t)NE * E * E, 0.f), beiv((size_t)NE * E, 0.f);
  for (int l = 0; l < L - 1; ++l) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "h_el_ion_%d/linear_0", l);
    const int kin = l == 0 ? 4 : E;
    const Leaf& w = leafp(buf, "w");
    const Leaf& b = leafp(buf, "b");
    for (int k = 0; k < kin; ++k)
      for (int f = 0; f < E; ++f) wei[((size_t)l * E + k) * E + f] = w.data[(size_t)k * E + f];
    for (int f = 0; f < E; ++f) beiv[(size_t)l * E + f] = b.data[f];
  }
  std::vector<float> whim((size_t)L * E * E, 0.f), bhimv((size_t)L * E, 0.f);
  for (int l = 0; l < L; ++l) {
    char buf[96];
    std::snprintf(buf, sizeof buf, "symm_features_%d/convolutional_features/h_ion_map/linear_0", l);
    const int g = l == 0 ? 4 : E;
    const Leaf& w = leafp(buf, "w");
    const Leaf& b = leafp(buf, "b");
    for (int k = 0; k < E; ++k)
      for (int f = 0; f < g; ++f) whim[((size_t)l * E + k) * E + f] = w.data[(size_t)k * g + f];
    for (int f = 0; f < g; ++f) bhimv[(size_t)l * E + f] = b.data[f];
  }
  dw_status st;
  if ((st = up(W.Wp, wp)) || (st = up(W.bp, bpv)) || (st = up(W.Wei, wei)) || (st = up(W.bei, beiv)) ||
      (st = up(W.Whim, whim)) || (st = up(W.bhim, bhimv)) || (st = up(W.table, leafp("input/h_ion", "embeddings").data)))
    return st;
  for (auto* v : {&W.WhmT, &W.WhmT_hi, &W.WhmT_lo, &W.bhm, &W.WhmB}) v->resize(L);
  for (int l = 0; l < L; ++l) {
    char buf[96];
    std::snprintf(buf, sizeof buf, "symm_features_%d/convolutional_features/h_map/linear_0", l);
    const Leaf& w = leafp(buf, "w");
    const int Kh = m->Kh[l], Khp = m->Kt0[l];
    std::vector<float> wt((size_t)128 * Khp, 0.f);
    for (int f = 0; f < E; ++f)
      for (int k = 0; k < Kh; ++k) wt[(size_t)f * Khp + k] = w.data[(size_t)k * E + f];
    std::vector<float> wb((size_t)pad_to(Kh, 128) * E, 0.f);
    std::copy(w.data.begin(), w.data.end(), wb.begin());
    if ((st = up(W.WhmT[l], wt)) || (st = up(W.bhm[l], leafp(buf, "b").data)) || (st = up(W.WhmB[l], wb))) return st;
    if (split && (st = upload_split(W.WhmT_hi[l], W.WhmT_lo[l], wt))) return st;
  }
  return DW_OK;
}

// Build the device layouts of all weights from the host leaves, in float or double.
dw_status prepare_weights(dw_model* m, Weights& W, bool dbl) {
  for (auto& e : m->expected)
    if (!m->leaves[e.first].set) return fail(DW_ERR_STATE, "parameter not set: " + e.first);
  const int W1 = m->W1, W2 = m->W2;
  const bool split = !dbl && m->prec != GEMM_F32;
  for (auto* v : {&W.Wmain, &W.Wmu, &W.bias, &W.Wmain_hi, &W.Wmain_lo, &W.Wmu_hi, &W.Wmu_lo, &W.WmainT, &W.WmuT})
    v->resize(m->L);
  auto up = [&](DevBuf& b, const std::vector<float>& v) { return dbl ? upload_as<double>(b, v) : upload(b, v); };
  for (int l = 0; l < m->L; ++l) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "h_el_%d/linear_0", l);
    const Leaf& w = m->leaves[leaf_key(buf, "w")];
    const Leaf& b = m->leaves[leaf_key(buf, "b")];
    const int Kh = m->Kh[l], Kx = m->Kx[l], Km = m->Km[l], Kt0 = m->Kt0[l], Kt = m->Kt[l];
    // reference row blocks of W: [h (Kh) | mean_up h (Kh) | mean_dn h (Kh) | 2-el block (Kt)], the 2-el block
    // being [pair up | pair dn] (FermiNet) or [e_ee | e_ei] (dpe4); X_l holds h in columns [0, Kh) and the
    // 2-el block in [Kt0, Kt0 + Kt)
    std::vector<float> wm((size_t)W1 * Kx, 0.f), wu((size_t)W1 * Km, 0.f);
    for (int o = 0; o < W1; ++o) {
      for (int k = 0; k < Kh; ++k) wm[(size_t)o * Kx + k] = w.data[(size_t)k * W1 + o];
      for (int k = 0; k < Kt; ++k) wm[(size_t)o * Kx + Kt0 + k] = w.data[(size_t)(3 * Kh + k) * W1 + o];
      for (int k = 0; k < 2 * Kh; ++k) wu[(size_t)o * Km + k] = w.data[(size_t)(Kh + k) * W1 + o];
    }
    // transposed copies [pad128(K)][W1] for the gradient pass (dX = dz W^T as C = A Bt^T)
    std::vector<float> wmT((size_t)pad_to(Kx, 128) * W1, 0.f), wuT((size_t)pad_to(Km, 128) * W1, 0.f);
    for (int o = 0; o < W1; ++o) {
      for (int k = 0; k < Kx; ++k) wmT[(size_t)k * W1 + o] = wm[(size_t)o * Kx + k];
      for (int k = 0; k < Km; ++k) wuT[(size_t)k * W1 + o] = wu[(size_t)o * Km + k];
    }
    dw_status st;
    if ((st = up(W.Wmain[l], wm))) return st;
    if ((st = up(W.Wmu[l], wu))) return st;
    if ((st = up(W.WmainT[l], wmT))) return st;
    if ((st = up(W.WmuT[l], wuT))) return st;
    if ((st = up(W.bias[l], b.data))) return st;
    if (split) {
      if ((st = upload_split(W.Wmain_hi[l], W.Wmain_lo[l], wm))) return st;
      if ((st = upload_split(W.Wmu_hi[l], W.Wmu_lo[l], wu))) return st;
    }
  }
  if (m->dpe4) {
    dw_status st = prepare_dpe4(m, W, dbl, split);
    if (st) return st;
  } else {
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
    if ((st = up(W.W2e, w2))) return st;
    if ((st = up(W.b2e, b2))) return st;
  }
  const char* bfk[2] = {"bf_up/linear_0", "bf_dn/linear_0"};
  const char* ak[2] = {"alpha_up", "alpha_dn"};
  const char* wk[2] = {"weights_up", "weights_dn"};
  std::vector<double> kap((size_t)2 * m->I * m->NO), wen((size_t)2 * m->I * m->NO);
  for (int s = 0; s < 2; ++s) {
    const Leaf& w = m->leaves[leaf_key(bfk[s], "w")];
    std::vector<float> wo((size_t)m->NOp * W1, 0.f);
    for (int c = 0; c < m->NO; ++c)
      for (int k = 0; k < W1; ++k) wo[(size_t)c * W1 + k] = w.data[(size_t)k * m->NO + c];
    std::vector<float> woT((size_t)pad_to(W1, 128) * m->NOk, 0.f);
    for (int c = 0; c < m->NO; ++c)
      for (int k = 0; k < W1; ++k) woT[(size_t)k * m->NOk + c] = w.data[(size_t)k * m->NO + c];
    dw_status st;
    if ((st = up(W.Wo[s], wo))) return st;
    if ((st = up(W.WoT[s], woT))) return st;
    if (split && (st = upload_split(W.Wo_hi[s], W.Wo_lo[s], wo))) return st;
    const Leaf& a = m->leaves[leaf_key("envelope_orbitals", ak[s])];
    const Leaf& we = m->leaves[leaf_key("envelope_orbitals", wk[s])];
    for (size_t q = 0; q < (size_t)m->I * m->NO; ++q) {
      const double x = a.data[q];
      kap[(size_t)s * m->I * m->NO + q] = x > 30.0 ? x : std::log1p(std::exp(x));  // softplus
      wen[(size_t)s * m->I * m->NO + q] = we.data[q];
    }
  }
  dw_status st;
  if (dbl) {
    if ((st = upload(W.kappa, kap))) return st;
    if ((st = upload(W.wenv, wen))) return st;
  } else {
    std::vector<float> kf(kap.begin(), kap.end()), wf(wen.begin(), wen.end());
    if ((st = upload(W.kappa, kf))) return st;
    if ((st = upload(W.wenv, wf))) return st;
  }
  // the uploads are hipMemcpy on the null stream, whose DMA may still be in flight on return: wait for the
  // null stream only (a device-wide synchronisation would also wait for other models' streams)
  HIP_TRY(hipStreamSynchronize(nullptr));
  W.ready = true;
  return DW_OK;
}

// Device weights for the model's precision (and the double layout when need_double).  A re-upload
// overwrites buffers that kernels queued earlier on `s` may still read, so `s` is drained first; the
// uploads themselves are complete when prepare_weights returns.
dw_status prepare(dw_model* m, bool need_double, hipStream_t s) {
  if (m->dirty) {
    if (m->wf.ready || m->wd.ready) HIP_TRY(hipStreamSynchronize(s));
    m->wf.ready = false;
    m->wd.ready = false;
    m->dirty = false;
  }
  const bool need_float = m->prec != GEMM_F64;
  if (need_float && !m->wf.ready) {
    dw_status st = prepare_weights(m, m->wf, false);
    if (st) {
      m->dirty = true;
      return st;
    }
  }
  if ((need_double || !need_float) && !m->wd.ready) {
    dw_status st = prepare_weights(m, m->wd, true);
    if (st) {
      m->dirty = true;
      return st;
    }
  }
  return DW_OK;
}

// Lay out the scratch of one pass of B walkers with C channels at `base` (nullptr: count only) and
// return its size in bytes.
template <typename T>
size_t layout_pass(const dw_model* m, long B, int C, char* base, PassBufs<T>* out) {
  const int Cp = rows_per_electron(C);
  const long rows = B * m->n * Cp;
  int kmmax = 0;
  for (int l = 0; l < m->L; ++l) kmmax = std::max(kmmax, m->Km[l]);
  Carver cv{base};
  PassBufs<T> b;
  for (int l = 0; l < m->L; ++l) b.X[l] = cv.take<T>((size_t)rows * m->Kx[l]);
  b.MU = cv.take<T>((size_t)B * Cp * kmmax);
  b.V = cv.take<T>((size_t)B * Cp * m->W1 * (C == 1 ? 4 : 1));  // value pass: up to 4 K-split parts
  if (sizeof(T) == 8 && C > 1) b.Z = cv.take<T>((size_t)rows * m->W1);  // unfused pre-activations (FP64 form)
  b.H = cv.take<T>((size_t)rows * m->W1);
  b.O = cv.take<T>((size_t)rows * m->NO);
  if (!det_reads_orbitals(m->n, C)) b.A = cv.take<T>((size_t)B * m->D * C * m->n * m->n);  // envelope x orbitals
  if (m->dpe4) {
    b.WM = cv.take<T>((size_t)m->L * B * m->n * m->n * (C > 1 ? 5 : 1) * 32);
    b.HM = cv.take<T>((size_t)rows * 32);
    b.HIM = cv.take<T>((size_t)m->L * m->I * 32);
  }
  b.ld = cv.take<double>((size_t)B * m->D);
  b.sd = cv.take<double>((size_t)B * m->D);
  b.dcond = cv.take<float>((size_t)B * m->D);
  b.g = cv.take<double>((size_t)std::max<long>(1, C > 1 ? B * m->D * 3 * m->n : 1));
  b.lap = cv.take<double>((size_t)std::max<long>(1, C > 1 ? B * m->D : 1));
  if (out) *out = b;
  return cv.off;
}

size_t pass_bytes(const dw_model* m, long B, int C, bool dbl) {
  return dbl ? layout_pass<double>(m, B, C, nullptr, nullptr) : layout_pass<float>(m, B, C, nullptr, nullptr);
}

// Scratch of one pass from the caller-owned workspace, else from the model's arena (grown on demand).
template <typename T>
dw_status pass_scratch(dw_model* m, long B, int C, PassBufs<T>* pb) {
  const size_t need = layout_pass<T>(m, B, C, nullptr, nullptr);
  char* base;
  if (m->ext_ws) {
    if (need > m->ext_ws_bytes)
      return fail(DW_ERR_INVALID, "the workspace given to dw_model_set_workspace is smaller than dw_workspace_size");
    base = (char*)m->ext_ws;
  } else {
    DevBuf& a = sizeof(T) == 8 ? m->arena_d : m->arena_f;
    HIP_TRY(a.ensure(need));
    base = a.as<char>();
  }
  layout_pass<T>(m, B, C, base, pb);
  return DW_OK;
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

// Bracket a region with events on stream s when timing is on (kind < 0: not timed).
struct Timed {
  dw_model* m;
  int kind;
  hipStream_t s;
  hipEvent_t a = nullptr;
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

// ---- per-type GEMM front ends ----
void gemm_t(const dw_model* m, const float* A, int lda, const float* Bt, const DevBuf& hi, const DevBuf& lo, int ldb,
            float* C, int ldc, long M, int N, int K, const GemmEpi& ep, hipStream_t s) {
  gemm(m->prec, A, lda, Bt, hi.as<uint16_t>(), lo.as<uint16_t>(), ldb, C, ldc, (int)M, N, K, ep, s);
}
void gemm_t(const dw_model*, const double* A, int lda, const double* Bt, const DevBuf&, const DevBuf&, int ldb,
            double* C, int ldc, long M, int N, int K, const GemmEpiD& ep, hipStream_t s) {
  gemm_f64(A, lda, Bt, ldb, C, ldc, (int)M, N, K, ep, s);
}

// Gaussian proposal of an MCMC step, formed inside a value pass: r' = r + stepsize * normals -> r_out.
struct Proposal {
  const float* normals;
  const void* chain;
  float* r_out;
};

// One pass of the wavefunction over B walkers with C channels (C == 1: value; C == 3n+2: Laplacian).
// `ws` is the pass's scratch (layout_pass); out_kappa (may be null) receives the per-walker conditioning.
template <typename T>
dw_status run_pass(dw_model* m, const PassBufs<T>& ws, const float* r, long B, const float* R, const float* Z, int C,
                   hipStream_t s, float* out_lp, float* out_phase, float* out_eloc, uint32_t* out_flags,
                   float* out_kappa, float cond_thr, const Proposal* prop = nullptr) {
  constexpr bool DBL = sizeof(T) == 8;
  Weights& W = DBL ? m->wd : m->wf;
  const int Cp = rows_per_electron(C);
  Shape sh{(int)B, m->n, m->u, m->I, m->D, C, Cp, m->W1, m->W2, m->L};
  const long rows = B * m->n * Cp;
  Timed pass_timer(m, C > 1 ? DW_TIMER_LOCAL_ENERGY : DW_TIMER_FORWARD, s);
  T* const* X = ws.X;
  if (prop) {  // the pass runs on the Gaussian proposal, formed and stored by the features kernel
    launch_features0<T>(r, R, X[0], m->Kx[0], m->Kt0[0], m->Kt[0], sh, s, prop->normals, prop->chain, prop->r_out);
    r = prop->r_out;
  } else {
    launch_features0<T>(r, R, X[0], m->Kx[0], m->Kt0[0], m->Kt[0], sh, s);
  }
  if (m->dpe4) {
    launch_dpe4_pairs<T>(r, W.Wp.as<T>(), W.bp.as<T>(), ws.WM, sh, s);
    launch_dpe4_elion<T>(r, R, Z, W.Wei.as<T>(), W.bei.as<T>(), W.Whim.as<T>(), W.bhim.as<T>(), W.table.as<T>(),
                         m->types, m->z_min, ws.HIM, X, m->Kx.data(), m->Kt0.data(), sh, s);
  } else {
    launch_twoel<T>(r, X, m->Kx.data(), m->Kh.data(), W.W2e.as<T>(), W.b2e.as<T>(), sh, s);
  }
  const long vstride = B * Cp * (long)m->W1;
  const long wm_layer = B * (long)m->n * m->n * (C > 1 ? 5 : 1) * 32;
  for (int l = 0; l < m->L; ++l) {
    if (m->dpe4) {  // e_ee of this iteration: hm = tanh rule of (X_l[:, h] W_map), then the convolution
      GemmEpiT<T> eh;
      eh.tile = 0;  // 64 x 64 tiles: the output is only 32 wide
      gemm_t(m, X[l], m->Kx[l], W.WhmT[l].as<T>(), W.WhmT_hi[l], W.WhmT_lo[l], m->Kt0[l], ws.HM, 32, rows, 32,
             m->Kt0[l], eh, s);
      launch_dpe4_tanh<T>(ws.HM, W.bhm[l].as<T>(), sh, s);
      launch_dpe4_conv<T>(ws.WM + l * wm_layer, ws.HM, X[l], m->Kx[l], m->Kt0[l], sh, s);
    }
    // Laplacian pass, FP32 form: spin means formed inside the per-walker GEMM (same values as below)
    bool fused_mu = false;
    if constexpr (!DBL) {
      if (C > 1 && m->prec == GEMM_F32)
        fused_mu = gemm_mean_lap(X[l], m->Kx[l], m->Kh[l], W.Wmu[l].as<float>(), m->Km[l], W.bias[l].as<float>(),
                                 ws.V, m->W1, sh, s);
    }
    if (!fused_mu) launch_spin_mean<T>(X[l], m->Kx[l], m->Kh[l], ws.MU, m->Km[l], sh, s);
    // value pass: split K of the per-walker spin-mean GEMM so that it fills the chip (its M is only B)
    int vs = 1;
    if (!DBL && C == 1 && m->prec == GEMM_F32) {
      if (m->tune.fwd_mu_split > 0) {
        vs = m->tune.fwd_mu_split;
      } else {
        // (2048 walkers x 256 columns = 128 tiles: a value pass 468 us with 4 parts, 455 us with 2)
        const long tiles64 = ((B * Cp + 63) / 64) * ((m->W1 + 63) / 64);
        vs = tiles64 >= 512 ? 1 : (tiles64 >= 128 ? 2 : 4);
      }
      while (vs > 1 && m->Km[l] % (16 * vs) != 0) vs >>= 1;  // each K slice a multiple of the 16-wide K tile
    }
    GemmEpiT<T> ev;
    ev.bias = W.bias[l].as<T>();
    ev.C = Cp;
    ev.nC = m->n * Cp;
    ev.ksplit = vs;
    ev.cstride = vstride;
    ev.tile = C == 1 ? -1 : m->tune.lap_mu_tile;
    if (!fused_mu)
      gemm_t(m, ws.MU, m->Km[l], W.Wmu[l].as<T>(), W.Wmu_hi[l], W.Wmu_lo[l], m->Km[l], ws.V, m->W1,
             B * Cp, m->W1, m->Km[l], ev, s);
    const bool last = (l == m->L - 1);
    T* outp = last ? ws.H : X[l + 1];
    const int ldo = last ? m->W1 : m->Kx[l + 1];
    GemmEpiT<T> em;
    em.V = ws.V;
    em.ldv = m->W1;
    em.C = Cp;
    em.nC = m->n * Cp;
    em.n = m->n;
    em.u = m->u;
    em.B = (int)B;
    em.vsplit = vs;
    em.vstride = vstride;
    if (C == 1) {
      em.tile = m->tune.fwd_main_tile;
      em.act_tanh = 1;
      em.permute = last ? 1 : 0;
      if constexpr (!DBL) {
        if (m->prec == GEMM_F32 && m->tune.fwd_main_tile == 3) {  // 16x16-MFMA tiles of 96 electrons
          gemm_value16(X[l], m->Kx[l], W.Wmain[l].as<float>(), m->Kx[l], ws.V, m->W1, vs, vstride, outp,
                       ldo, last ? 1 : 0, sh, s);
          continue;
        }
      }
      gemm_t(m, X[l], m->Kx[l], W.Wmain[l].as<T>(), W.Wmain_hi[l], W.Wmain_lo[l], m->Kx[l], outp, ldo, rows, m->W1,
             m->Kx[l], em, s);
    } else {
      Timed t(m, l >= 1 ? DW_TIMER_GEMM_1EL_LAP : -1, s);
      if constexpr (DBL) {
        gemm_t(m, X[l], m->Kx[l], W.Wmain[l].as<T>(), W.Wmain_hi[l], W.Wmain_lo[l], m->Kx[l], ws.Z, m->W1,
               rows, m->W1, m->Kx[l], em, s);
        launch_tanh_rule<T>(ws.Z, outp, ldo, last ? 1 : 0, sh, s);
      } else {
        gemm_lap(m->prec, X[l], m->Kx[l], W.Wmain[l].as<float>(), W.Wmain_hi[l].as<uint16_t>(),
                 W.Wmain_lo[l].as<uint16_t>(), m->Kx[l], ws.V, m->W1, outp, ldo, last ? 1 : 0, sh,
                 m->tune.lap_electrons, m->tune.lap_cols, s);
      }
    }
  }
  // orbital GEMMs on the spin-major rows
  const long up_rows = B * m->u * Cp, dn_rows = B * m->dn * Cp;
  GemmEpiT<T> e0;
  e0.tile = C == 1 ? m->tune.fwd_orb_tile : m->tune.lap_orb_tile;
  gemm_t(m, ws.H, m->W1, W.Wo[0].as<T>(), W.Wo_hi[0], W.Wo_lo[0], m->W1, ws.O, m->NO, up_rows, m->NO,
         m->W1, e0, s);
  gemm_t(m, ws.H + up_rows * m->W1, m->W1, W.Wo[1].as<T>(), W.Wo_hi[1], W.Wo_lo[1], m->W1,
         ws.O + up_rows * m->NO, m->NO, dn_rows, m->NO, m->W1, e0, s);
  const OrbSrc<T> src{ws.O, m->NO, r, R, W.kappa.as<T>(), W.wenv.as<T>()};
  if (!det_reads_orbitals(m->n, C))
    launch_envelope<T>(ws.O, m->NO, r, R, W.kappa.as<T>(), W.wenv.as<T>(), ws.A, sh, s);
  bool combined = false;
  {
    Timed t(m, C > 1 ? DW_TIMER_DET_LAP : -1, s);
    const CombineValue cv{out_lp, out_phase, out_flags, out_kappa, cond_thr, 0};
    combined = launch_det<T>(ws.A, src, ws.ld, ws.sd, ws.g, ws.lap, ws.dcond, sh, s, C == 1 ? &cv : nullptr);
  }
  if (!combined)
    launch_combine(ws.ld, ws.sd, ws.g, ws.lap, ws.dcond, cond_thr, r, R, Z, out_lp, out_phase, out_eloc, out_flags,
                   out_kappa, sh, s);
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

// Offset (floats) of a parameter leaf in the packed gradient: the leaves in the order of m->expected.
long leaf_offset(const dw_model* m, const std::string& key) {
  long off = 0;
  for (auto& e : m->expected) {
    if (e.first == key) return off;
    long cnt = 1;
    for (auto v : e.second) cnt *= v;
    off += cnt;
  }
  return -1;
}

// Scratch of the gradient pass (after the value pass's own scratch).
template <typename T>
struct GradBufs {
  double* gl = nullptr;
  T *dbf = nullptr, *denv = nullptr, *dH = nullptr, *dz = nullptr, *DZS = nullptr, *dMU = nullptr, *MU = nullptr;
  T* dX[4] = {nullptr, nullptr, nullptr, nullptr};
  T* X2[3] = {nullptr, nullptr, nullptr};
  T* DP[3] = {nullptr, nullptr, nullptr};
  T *part = nullptr, *cpart = nullptr, *epart = nullptr;
  // dpe4: h_map rows per layer, their gradient, the dh product, pair-map gradients (natural pair order), pair and
  // el-ion stream rows, the el-ion products summed into dhim (float)
  T *HMg[4] = {}, *dhm = nullptr, *tmp = nullptr, *DWp[4] = {}, *X2p[4] = {}, *DZm[4] = {}, *DZh[4] = {};
  T *Q[4] = {}, *DZQ[4] = {}, *PH[4] = {};
  float* dhim = nullptr;
};

constexpr int kEnvWalkersPerChunk = 32;

template <typename T>
size_t layout_grad(const dw_model* m, long B, char* base, GradBufs<T>* out) {
  const long rows = B * m->n, pairs = B * m->n * m->n;
  int kmmax = 0;
  for (int l = 0; l < m->L; ++l) kmmax = std::max(kmmax, m->Km[l]);
  Carver cv{base};
  GradBufs<T> g;
  g.gl = cv.take<double>((size_t)B * m->D);
  g.dbf = cv.take<T>((size_t)rows * m->NOk);
  g.denv = cv.take<T>((size_t)rows * m->NOk);
  g.dH = cv.take<T>((size_t)rows * m->W1);
  g.dz = cv.take<T>((size_t)rows * m->W1);
  g.DZS = cv.take<T>((size_t)B * m->W1);
  g.dMU = cv.take<T>((size_t)B * kmmax);
  g.MU = cv.take<T>((size_t)B * kmmax);
  for (int l = m->dpe4 ? 0 : 1; l < m->L; ++l) g.dX[l] = cv.take<T>((size_t)rows * m->Kx[l]);
  if (!m->dpe4)
    for (int l = 0; l < m->L - 1; ++l) {
      g.X2[l] = cv.take<T>((size_t)pairs * (l == 0 ? 4 : 32));
      g.DP[l] = cv.take<T>((size_t)pairs * 32);
    }
  const long eirows = rows * m->I;
  if (m->dpe4) {
    for (int l = 0; l < m->L; ++l) {
      g.HMg[l] = cv.take<T>((size_t)rows * 32);
      g.DWp[l] = cv.take<T>((size_t)pairs * 32);
      g.X2p[l] = cv.take<T>((size_t)pairs * 32);
      g.DZm[l] = cv.take<T>((size_t)pairs * 32);
      if (l < m->L - 1) g.DZh[l] = cv.take<T>((size_t)pairs * 32);
      g.Q[l] = cv.take<T>((size_t)eirows * 32);
      if (l < m->L - 1) g.DZQ[l] = cv.take<T>((size_t)eirows * 32);
      g.PH[l] = cv.take<T>((size_t)eirows * 32);
    }
    g.dhm = cv.take<T>((size_t)rows * 32);
    g.tmp = cv.take<T>((size_t)rows * m->W1);
    g.dhim = cv.take<float>((size_t)m->L * m->I * 32);
  }
  size_t pmax = 0;
  for (int l = 0; l < m->L; ++l) {
    pmax = std::max(pmax, wgrad_part_elems(m->Kh[l], m->W1, rows));
    pmax = std::max(pmax, wgrad_part_elems(2 * m->Ke[l], m->W1, rows));
    pmax = std::max(pmax, wgrad_part_elems(2 * m->Kh[l], m->W1, B));
  }
  pmax = std::max(pmax, wgrad_part_elems(32, 32, pairs));
  pmax = std::max(pmax, wgrad_part_elems(m->W1, m->NO, B * std::max(m->u, m->dn)));
  if (m->dpe4) {
    for (int l = 0; l < m->L; ++l) pmax = std::max(pmax, wgrad_part_elems(m->Kh[l], 32, rows));
    pmax = std::max(pmax, wgrad_part_elems(32, 32, eirows));
  }
  g.part = cv.take<T>(pmax);
  size_t cmax = (size_t)std::max(colsum_chunks(pairs), colsum_chunks(B)) * std::max(32, m->W1);
  if (m->dpe4)
    cmax = std::max({cmax, (size_t)colsum_chunks(rows) * m->I * 32, (size_t)colsum_chunks(eirows) * 32,
                     (size_t)colsum_chunks(rows) * 32});
  g.cpart = cv.take<T>(cmax);
  g.epart = cv.take<T>((size_t)((B + kEnvWalkersPerChunk - 1) / kEnvWalkersPerChunk) * 4 * m->I * m->NO);
  if (out) *out = g;
  return cv.off;
}

// Parameter gradient G = sum_b c_b d log_psi_sqr(r_b) / d theta into the packed float buffer `grad`.
template <typename T>
dw_status run_grad(dw_model* m, const float* r, long B, const float* R, const float* Z, const float* c, float* grad,
                   float* out_lp, hipStream_t s) {
  constexpr bool DBL = sizeof(T) == 8;
  Weights& W = DBL ? m->wd : m->wf;
  PassBufs<T> pb;
  dw_status st = pass_scratch(m, B, 1, &pb);
  if (st) return st;
  st = run_pass<T>(m, pb, r, B, R, Z, 1, s, out_lp, nullptr, nullptr, nullptr, nullptr, 0.f);
  if (st) return st;
  HIP_TRY(m->grad_arena.ensure(layout_grad<T>(m, B, nullptr, nullptr)));
  GradBufs<T> g;
  layout_grad<T>(m, B, m->grad_arena.as<char>(), &g);
  const Shape sh{(int)B, m->n, m->u, m->I, m->D, 1, 1, m->W1, m->W2, m->L};
  const long rows = B * m->n, pairs = rows * m->n;
  const long up_rows = B * m->u, dn_rows = B * m->dn;
  auto mm = [&](const T* A, int lda, const T* Bt, int ldb, T* C, int ldc, long M, int N, int K) {
    if constexpr (DBL) {
      GemmEpiD e;
      gemm_f64(A, lda, Bt, ldb, C, ldc, (int)M, N, K, e, s);
    } else {
      GemmEpi e;
      gemm(GEMM_F32, A, lda, Bt, nullptr, nullptr, ldb, C, ldc, (int)M, N, K, e, s);
    }
  };
  auto off = [&](const char* mod, const char* leaf) { return grad + leaf_offset(m, leaf_key(mod, leaf)); };
  // determinants -> orbitals and envelope
  launch_grad_seed(pb.ld, pb.sd, c, g.gl, sh, s);
  const OrbSrc<T> src{pb.O, m->NO, r, R, W.kappa.as<T>(), W.wenv.as<T>()};
  launch_det_bwd<T>(src, g.gl, g.dbf, g.denv, m->NOk, sh, s);
  {
    const int chunks = launch_env_grad<T>(g.denv, m->NOk, r, R, W.kappa.as<T>(), W.wenv.as<T>(), g.epart,
                                          kEnvWalkersPerChunk, sh, s);
    launch_sum_parts<T>(g.epart, chunks, 4L * m->I * m->NO, off("envelope_orbitals", "alpha_up"), s);
  }
  const char* bfk[2] = {"bf_up/linear_0", "bf_dn/linear_0"};
  for (int sp = 0; sp < 2; ++sp) {
    const long r0 = sp ? up_rows : 0, nr = sp ? dn_rows : up_rows;
    wgrad<T>(pb.H + r0 * m->W1, m->W1, g.dbf + r0 * m->NOk, m->NOk, m->W1, m->NO, nr, g.part, off(bfk[sp], "w"), m->NO,
             s);
    mm(g.dbf + r0 * m->NOk, m->NOk, W.WoT[sp].as<T>(), m->NOk, g.dH + r0 * m->W1, m->W1, nr, m->W1, m->NOk);
  }
  // dpe4: hm_l of every layer (the value pass keeps only the last one)
  const long wm_layer = B * (long)m->n * m->n * 32;
  if (m->dpe4)
    for (int l = 0; l < m->L; ++l) {
      mm(pb.X[l], m->Kx[l], W.WhmT[l].as<T>(), m->Kt0[l], g.HMg[l], 32, rows, 32, m->Kt0[l]);
      launch_dpe4_tanh<T>(g.HMg[l], W.bhm[l].as<T>(), sh, s);
    }
  // 1-el layers, last to first
  launch_dz_last<T>(g.dH, pb.H, g.dz, sh, s);
  for (int l = m->L - 1; l >= 0; --l) {
    char key[64];
    std::snprintf(key, sizeof key, "h_el_%d/linear_0", l);
    const int Kh = m->Kh[l], Kx = m->Kx[l], Km = m->Km[l];
    float* gw = off(key, "w");
    launch_walker_sum<T>(g.dz, m->W1, m->W1, g.DZS, sh, s);
    launch_spin_mean<T>(pb.X[l], Kx, Kh, g.MU, Km, sh, s);
    wgrad<T>(pb.X[l], Kx, g.dz, m->W1, Kh, m->W1, rows, g.part, gw, m->W1, s);
    wgrad<T>(g.MU, Km, g.DZS, m->W1, 2 * Kh, m->W1, B, g.part, gw + (long)Kh * m->W1, m->W1, s);
    wgrad<T>(pb.X[l] + m->Kt0[l], Kx, g.dz, m->W1, m->Kt[l], m->W1, rows, g.part, gw + 3L * Kh * m->W1, m->W1, s);
    launch_colsum<T>(g.DZS, m->W1, B, m->W1, g.cpart, off(key, "b"), s);
    if (l == 0 && !m->dpe4) break;
    mm(g.dz, m->W1, W.WmainT[l].as<T>(), m->W1, g.dX[l], Kx, rows, Kx, m->W1);
    if (m->dpe4) {  // e_ee_l = conv(w_l, hm_l), hm_l = tanh(h_l W_map + b): gradients of W_map, b, w_l and h_l
      char hk[96];
      std::snprintf(hk, sizeof hk, "symm_features_%d/convolutional_features/h_map/linear_0", l);
      launch_dpe4_conv_bwd<T>(g.dX[l], Kx, m->Kt0[l], pb.WM + l * wm_layer, g.HMg[l], g.dhm, g.DWp[l], sh, s);
      launch_dtanh<T>(g.dhm, g.HMg[l], rows * 32, s);
      wgrad<T>(pb.X[l], Kx, g.dhm, 32, Kh, 32, rows, g.part, off(hk, "w"), 32, s);
      launch_colsum<T>(g.dhm, 32, rows, 32, g.cpart, off(hk, "b"), s);
      if (l > 0) {
        mm(g.dhm, 32, W.WhmB[l].as<T>(), 32, g.tmp, Kh, rows, Kh, 32);
        launch_add_cols<T>(g.dX[l], Kx, g.tmp, Kh, rows, Kh, s);
      }
    }
    if (l == 0) break;
    mm(g.DZS, m->W1, W.WmuT[l].as<T>(), m->W1, g.dMU, Km, B, Km, m->W1);
    launch_dz<T>(g.dX[l], Kx, g.dMU, Km, pb.X[l], Kx, Kh, g.dz, sh, s);
  }
  if (m->dpe4) {
    // pair streams: rows grouped by spin relation (same-spin pairs of every walker, then diff-spin pairs)
    T* wml[4] = {};
    for (int l = 0; l < m->L; ++l) wml[l] = pb.WM + l * wm_layer;
    launch_dpe4_pairs_bwd<T>(r, W.Wp.as<T>(), W.bp.as<T>(), g.DWp, wml, g.X2p, g.DZm, g.DZh, sh, s);
    const long P0 = B * ((long)m->u * m->u + (long)m->dn * m->dn), P1 = B * 2L * m->u * m->dn;
    for (int l = 0; l < m->L; ++l) {
      const int kin = l == 0 ? 1 : 32;
      for (int sp = 0; sp < 2; ++sp) {
        const long r0 = sp ? P0 : 0, nr = sp ? P1 : P0;
        if (nr == 0) continue;
        char key[96];
        std::snprintf(key, sizeof key, "symm_features_%d/convolutional_features/%s/linear_0", l,
                      sp ? "w_diff" : "w_same");
        wgrad<T>(g.X2p[l] + r0 * 32, 32, g.DZm[l] + r0 * 32, 32, kin, 32, nr, g.part, off(key, "w"), 32, s);
        launch_colsum<T>(g.DZm[l] + r0 * 32, 32, nr, 32, g.cpart, off(key, "b"), s);
        if (l < m->L - 1) {
          std::snprintf(key, sizeof key, "%s_%d/linear_0", sp ? "h_diff" : "h_same", l);
          wgrad<T>(g.X2p[l] + r0 * 32, 32, g.DZh[l] + r0 * 32, 32, kin, 32, nr, g.part, off(key, "w"), 32, s);
          launch_colsum<T>(g.DZh[l] + r0 * 32, 32, nr, 32, g.cpart, off(key, "b"), s);
        }
      }
    }
    // el-ion stream and the ion maps
    int ldx[4] = {}, col[4] = {};
    for (int l = 0; l < m->L; ++l) {
      ldx[l] = m->Kx[l];
      col[l] = m->Kt0[l] + 32;
    }
    launch_dpe4_elion_bwd<T>(r, R, W.Wei.as<T>(), W.bei.as<T>(), pb.HIM, g.dX, ldx, col, g.Q, g.DZQ, g.PH, sh, s);
    const long eirows = rows * m->I;
    float* gW[4] = {};
    float* gB[4] = {};
    for (int l = 0; l < m->L; ++l) {
      char key[96];
      if (l < m->L - 1) {
        std::snprintf(key, sizeof key, "h_el_ion_%d/linear_0", l);
        wgrad<T>(g.Q[l], 32, g.DZQ[l], 32, l == 0 ? 4 : 32, 32, eirows, g.part, off(key, "w"), 32, s);
        launch_colsum<T>(g.DZQ[l], 32, eirows, 32, g.cpart, off(key, "b"), s);
      }
      launch_colsum<T>(g.PH[l], m->I * 32, rows, m->I * 32, g.cpart, g.dhim + (long)l * m->I * 32, s);
      std::snprintf(key, sizeof key, "symm_features_%d/convolutional_features/h_ion_map/linear_0", l);
      gW[l] = off(key, "w");
      gB[l] = off(key, "b");
    }
    launch_dpe4_him_bwd<T>(g.dhim, pb.HIM, W.table.as<T>(), W.Whim.as<T>(), Z, m->z_min, m->types, m->L, m->I, gW,
                           gB, off("input/h_ion", "embeddings"), s);
  } else if (m->L > 1) {  // 2-el stream
    launch_twoel_bwd<T>(r, g.dX, m->Kx.data(), m->Kh.data(), W.W2e.as<T>(), W.b2e.as<T>(), g.X2, g.DP, sh, s);
    for (int l = 0; l < m->L - 1; ++l) {
      char key[64];
      std::snprintf(key, sizeof key, "h_el_el_%d/linear_0", l);
      const int kin = l == 0 ? 4 : 32;
      wgrad<T>(g.X2[l], kin, g.DP[l], 32, kin, 32, pairs, g.part, off(key, "w"), 32, s);
      launch_colsum<T>(g.DP[l], 32, pairs, 32, g.cpart, off(key, "b"), s);
    }
  }
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

// One pass in the model's precision, its scratch from the model (arena or caller-owned workspace).
dw_status pass(dw_model* m, const float* r, long B, const float* R, const float* Z, int C, hipStream_t s,
               float* out_lp, float* out_phase, float* out_eloc, uint32_t* out_flags, float* out_kappa,
               float cond_thr) {
  if (m->prec == GEMM_F64) {
    PassBufs<double> pb;
    dw_status st = pass_scratch(m, B, C, &pb);
    if (st) return st;
    return run_pass<double>(m, pb, r, B, R, Z, C, s, out_lp, out_phase, out_eloc, out_flags, out_kappa, 0.f);
  }
  PassBufs<float> pb;
  dw_status st = pass_scratch(m, B, C, &pb);
  if (st) return st;
  return run_pass<float>(m, pb, r, B, R, Z, C, s, out_lp, out_phase, out_eloc, out_flags, out_kappa, cond_thr);
}

}  // namespace

namespace dwmi {

dw_status set_error(dw_status code, const std::string& msg) { return fail(code, msg); }

dw_status mcmc_run(dw_model* m, const McmcArgs& a, const FixedScratch* fx, hipStream_t s) {
  if (!m || !a.r || !a.R || !a.Z || !a.log_psi_sqr || !a.walker_age || !a.chain)
    return fail(DW_ERR_INVALID, "null argument");
  if (a.n_walkers <= 0 || a.n_steps < 0) return fail(DW_ERR_INVALID, "bad sizes");
  if (a.n_steps > 0 && (!a.normals || !a.uniforms)) return fail(DW_ERR_INVALID, "draws required");
  const bool dbl = m->prec == GEMM_F64;
  const long B = a.n_walkers;
  const int cnt3 = (int)(B * m->n * 3);
  const int kind = m->prop_kind;
  const bool scaled = kind == DW_PROPOSAL_LOCAL || kind == DW_PROPOSAL_LANGEVIN || kind == DW_PROPOSAL_LOCAL_ONE_EL;
  const bool one_el = kind == DW_PROPOSAL_NORMAL_ONE_EL || kind == DW_PROPOSAL_LOCAL_ONE_EL;
  const size_t draw_stride = one_el ? (size_t)B * 3 : (size_t)cnt3;  // floats of draws per step
  float *rprop, *lpprop, *logq = nullptr;
  int32_t* cnt;
  PassBufs<float> pf;
  PassBufs<double> pd;
  if (fx) {
    char* const base = fx->check_only ? nullptr : fx->base;
    const size_t scratch_bytes = fx->bytes;
    Carver cv{base, pass_bytes(m, B, 1, dbl)};  // the pass's layout first, then the chain's own buffers
    rprop = cv.take<float>((size_t)cnt3);
    lpprop = cv.take<float>((size_t)B);
    if (scaled) logq = cv.take<float>((size_t)B);
    cnt = cv.take<int32_t>((size_t)2 * std::max(a.n_steps, 1));  // accepted counts, then tickets
    if (cv.off > scratch_bytes) {
      const size_t per_walker = (pass_bytes(m, 64, 1, dbl) + 64 * (size_t)(m->n * 3 + 2) * 4) / 64;
      return fail(DW_ERR_INVALID, "dw_mcmc_steps_noalloc: " + std::to_string(B) + " walkers need " +
                                      std::to_string(cv.off) + " bytes of scratch, more than its " +
                                      std::to_string(scratch_bytes) + " (about " +
                                      std::to_string(scratch_bytes / std::max<size_t>(per_walker, 1)) +
                                      " walkers of this model fit)");
    }
    if (m->dirty || !(dbl ? m->wd.ready : m->wf.ready))
      return fail(DW_ERR_STATE, "dw_mcmc_steps_noalloc: the weights are not on the device; call dw_logpsi, "
                                "dw_mcmc_steps or dw_local_energy once after the last dw_model_set_param");
    if (fx->check_only) return DW_OK;
    if (dbl)
      layout_pass<double>(m, B, 1, base, &pd);
    else
      layout_pass<float>(m, B, 1, base, &pf);
  } else {
    dw_status st = prepare(m, false, s);
    if (st) return st;
    HIP_TRY(m->rprop.ensure((size_t)cnt3 * 4));
    HIP_TRY(m->lpprop.ensure((size_t)B * 4));
    HIP_TRY(m->cnt.ensure((size_t)2 * std::max(a.n_steps, 1) * 4));  // accepted counts, then tickets
    rprop = m->rprop.as<float>();
    lpprop = m->lpprop.as<float>();
    cnt = m->cnt.as<int32_t>();
    if (scaled) {
      HIP_TRY(m->logq.ensure((size_t)B * 4));
      logq = m->logq.as<float>();
    }
    st = dbl ? pass_scratch(m, B, 1, &pd) : pass_scratch(m, B, 1, &pf);
    if (st) return st;
  }
  // the timing hooks create events; the non-allocating variant records none
  const bool timing = m->timing;
  if (fx) m->timing = false;
  auto value_pass = [&](const float* rin, float* lp_out, const Proposal* prop) {
    if (dbl)
      return run_pass<double>(m, pd, rin, B, a.R, a.Z, 1, s, lp_out, nullptr, nullptr, nullptr, nullptr, 0.f, prop);
    return run_pass<float>(m, pf, rin, B, a.R, a.Z, 1, s, lp_out, nullptr, nullptr, nullptr, nullptr, 0.f, prop);
  };
  // log_psi_sqr is re-evaluated on entry (mcmc.py:396)
  dw_status st = value_pass(a.r, a.log_psi_sqr, nullptr);
  if (!st && a.n_steps > 0) {
    launch_fill_i32(cnt, 0, 2 * a.n_steps, s);
    for (int k = 0; k < a.n_steps && !st; ++k) {
      const float* xi = a.normals + (size_t)k * draw_stride;
      if (one_el)
        launch_propose_one_el(kind, a.r, xi, a.chain, a.R, m->I, m->prop_rmin, m->prop_rmax, rprop, logq, (int)B,
                              m->n, s);
      else if (scaled)
        launch_propose_scaled(kind, a.r, xi, a.chain, a.R, a.Z, m->I, m->prop_rmin, m->prop_rmax, m->prop_scale, rprop,
                              logq, (int)B, m->n, s);
      // the Gaussian proposal is formed by the pass's first kernel, which stores it to rprop
      const Proposal gauss{xi, a.chain, rprop};
      st = (one_el || scaled) ? value_pass(rprop, lpprop, nullptr) : value_pass(a.r, lpprop, &gauss);
      if (st) break;
      launch_accept(a.r, rprop, a.log_psi_sqr, lpprop, logq, a.uniforms + (size_t)k * B, a.walker_age, a.max_age,
                    cnt + k, a.out_flags, (int)B, m->n, a.chain, cnt + a.n_steps + k,
                    a.out_acc ? a.out_acc + k : nullptr, a.stepsize_update_interval, a.target_acceptance,
                    a.min_stepsize, a.max_stepsize, s);
    }
  }
  m->timing = timing;
  if (st) return st;
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

}  // namespace dwmi

extern "C" {

const char* dw_last_error(void) { return g_err.c_str(); }
const char* dw_version(void) { return "dwmi 0.2 (gfx950)"; }

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
  if (d.n_el < 2 || d.n_el > 63) return fail(DW_ERR_INVALID, "n_el must be in [2, 63]");
  if (d.n_up < 1 || d.n_up >= d.n_el) return fail(DW_ERR_INVALID, "need 1 <= n_up < n_el (both spin blocks non-empty)");
  if (d.n_ion < 1 || d.n_ion > 64) return fail(DW_ERR_INVALID, "n_ion must be in [1, 64]");
  if (d.n_det < 1) return fail(DW_ERR_INVALID, "n_det must be >= 1");
  if (d.width2 != 32) return fail(DW_ERR_INVALID, "the fused 2-el kernel requires width2 == 32");
  if (d.width1 < 128 || d.width1 % 128 != 0) return fail(DW_ERR_INVALID, "width1 must be a multiple of 128");
  if (d.n_iter < 2 || d.n_iter > 4) return fail(DW_ERR_INVALID, "n_iter must be in [2, 4]");
  if (d.precision < DW_FP32 || d.precision > DW_FP64) return fail(DW_ERR_INVALID, "unknown precision");
  if (d.embedding != DW_EMBEDDING_FERMINET && d.embedding != DW_EMBEDDING_DPE4)
    return fail(DW_ERR_INVALID, "unknown embedding");
  if (d.embedding == DW_EMBEDDING_DPE4 && (d.n_ion_types < 1 || d.n_ion_types > 128))
    return fail(DW_ERR_INVALID, "dpe4 needs 1 <= n_ion_types <= 128");
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
  m->NOk = pad_to(m->NO, 32);
  m->prec = (GemmPrec)d.precision;
  m->dpe4 = d.embedding == DW_EMBEDDING_DPE4;
  m->z_min = d.z_min;
  m->types = d.n_ion_types;
  for (int l = 0; l < m->L; ++l) {
    const int Kh = (l == 0) ? 4 * m->I : m->W1;
    const int Ke = (l == 0) ? 4 : m->W2;
    m->Kh.push_back(Kh);
    m->Ke.push_back(Ke);
    // dpe4: X_l = [h | zeros to a multiple of 32 | e_ee (32) | e_ei (4 at l = 0, else 32)], so that the
    // h_map GEMM reads the h columns with K a multiple of 32
    const int t0 = m->dpe4 ? pad_to(Kh, 32) : Kh;
    const int tw = m->dpe4 ? 32 + (l == 0 ? 4 : 32) : 2 * Ke;
    m->Kt0.push_back(t0);
    m->Kt.push_back(tw);
    m->Kx.push_back(pad_to(t0 + tw, 32));
    m->Km.push_back(pad_to(2 * Kh, 32));
  }
  for (int l = 0; l < m->L; ++l) {
    char buf[96];
    std::snprintf(buf, sizeof buf, "h_el_%d/linear_0", l);
    m->expected.push_back({leaf_key(buf, "w"), {3 * m->Kh[l] + m->Kt[l], m->W1}});
    m->expected.push_back({leaf_key(buf, "b"), {m->W1}});
    if (m->dpe4) {
      const int fin = l == 0 ? 1 : 32, g = l == 0 ? 4 : 32;
      const struct { const char* mod; int64_t kin, kout; } conv[4] = {
          {"w_same", fin, 32}, {"w_diff", fin, 32}, {"h_map", m->Kh[l], 32}, {"h_ion_map", 32, g}};
      for (auto& c : conv) {
        std::snprintf(buf, sizeof buf, "symm_features_%d/convolutional_features/%s/linear_0", l, c.mod);
        m->expected.push_back({leaf_key(buf, "w"), {c.kin, c.kout}});
        m->expected.push_back({leaf_key(buf, "b"), {c.kout}});
      }
      if (l < m->L - 1) {
        const struct { const char* mod; int64_t kin; } strm[3] = {{"h_same", fin}, {"h_diff", fin}, {"h_el_ion", g}};
        for (auto& c : strm) {
          std::snprintf(buf, sizeof buf, "%s_%d/linear_0", c.mod, l);
          m->expected.push_back({leaf_key(buf, "w"), {c.kin, 32}});
          m->expected.push_back({leaf_key(buf, "b"), {32}});
        }
      }
    } else if (l < m->L - 1) {
      std::snprintf(buf, sizeof buf, "h_el_el_%d/linear_0", l);
      m->expected.push_back({leaf_key(buf, "w"), {l == 0 ? 4 : m->W2, m->W2}});
      m->expected.push_back({leaf_key(buf, "b"), {m->W2}});
    }
  }
  if (m->dpe4) m->expected.push_back({leaf_key("input/h_ion", "embeddings"), {m->types, 32}});
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

dw_status dw_model_set_tuning(dw_model* m, const dw_tuning* t) {
  if (!m || !t) return fail(DW_ERR_INVALID, "null argument");
  for (int32_t v : {t->fwd_orb_tile, t->lap_orb_tile, t->lap_mu_tile})
    if (v < -1 || v > 2) return fail(DW_ERR_INVALID, "tile must be -1, 0, 1 or 2");
  if (t->fwd_main_tile < -1 || t->fwd_main_tile > 4) return fail(DW_ERR_INVALID, "fwd_main_tile must be -1 .. 4");
  if (!(t->fwd_mu_split == 0 || t->fwd_mu_split == 1 || t->fwd_mu_split == 2 || t->fwd_mu_split == 4))
    return fail(DW_ERR_INVALID, "fwd_mu_split must be 0 (auto), 1, 2 or 4");
  const int Cp = rows_per_electron(3 * m->n + 2);
  if (t->lap_electrons < 0 || t->lap_electrons * Cp > 192)
    return fail(DW_ERR_INVALID, "lap_electrons * pad16(3 n + 2) must be <= 192");
  if (!(t->lap_cols == 0 || t->lap_cols == 128 || t->lap_cols == 256))
    return fail(DW_ERR_INVALID, "lap_cols must be 0, 128 or 256");
  m->tune = *t;
  return DW_OK;
}

dw_status dw_model_get_tuning(const dw_model* m, dw_tuning* out) {
  if (!m || !out) return fail(DW_ERR_INVALID, "null argument");
  *out = m->tune;
  return DW_OK;
}

dw_status dw_model_set_workspace(dw_model* m, void* device_ptr, size_t bytes) {
  if (!m) return fail(DW_ERR_INVALID, "null model");
  if (device_ptr && ((uintptr_t)device_ptr & 255)) return fail(DW_ERR_INVALID, "workspace must be 256-byte aligned");
  m->arena_f.release();
  m->arena_d.release();
  m->ext_ws = device_ptr;
  m->ext_ws_bytes = device_ptr ? bytes : 0;
  return DW_OK;
}

dw_status dw_model_set_chunk(dw_model* m, int32_t max_walkers_per_pass) {
  if (!m || max_walkers_per_pass < 0) return fail(DW_ERR_INVALID, "bad chunk");
  m->chunk = max_walkers_per_pass;
  return DW_OK;
}

size_t dw_workspace_size(const dw_model* m, int32_t n_walkers, int32_t mode) {
  if (!m || n_walkers <= 0) return 0;
  const bool dbl = m->prec == GEMM_F64;
  if (!mode) return pass_bytes(m, n_walkers, 1, dbl);
  const int C = 3 * m->n + 2;
  long B = n_walkers;
  if (m->chunk > 0 && B > m->chunk) B = m->chunk;
  size_t bytes = pass_bytes(m, B, C, dbl);
  // the FP64 recompute of flagged walkers carves double activations from the same region
  if (!dbl && m->d.fp64_cond_threshold > 0.f) bytes = std::max(bytes, pass_bytes(m, B, C, true));
  return bytes;
}

dw_status dw_model_conditioning(dw_model* m, float* out, int32_t n_walkers, dw_stream stream) {
  if (!m || !out || n_walkers <= 0) return fail(DW_ERR_INVALID, "bad argument");
  if (n_walkers > m->kappa_n)
    return fail(DW_ERR_STATE, "the last dw_logpsi / dw_local_energy call covered " + std::to_string(m->kappa_n) +
                                  " walkers");
  HIP_TRY(hipMemcpyAsync(out, m->kappa_w.p, (size_t)n_walkers * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return DW_OK;
}

dw_status dw_logpsi(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                    float* out_log_psi_sqr, float* out_phase, dw_stream stream) {
  if (!m || !r || !R || !Z || !out_log_psi_sqr) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  hipStream_t s = (hipStream_t)stream;
  dw_status st = prepare(m, false, s);
  if (st) return st;
  HIP_TRY(m->kappa_w.ensure((size_t)n_walkers * 4));
  m->kappa_n = n_walkers;
  return pass(m, r, n_walkers, R, Z, 1, s, out_log_psi_sqr, out_phase, nullptr, nullptr, m->kappa_w.as<float>(), 0.f);
}

dw_status dw_local_energy(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                          float* out_eloc, float* out_log_psi_sqr, uint32_t* out_flags, dw_stream stream) {
  if (!m || !r || !R || !Z || !out_eloc) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  const bool recompute = m->prec != GEMM_F64 && m->d.fp64_cond_threshold > 0.f;
  hipStream_t s = (hipStream_t)stream;
  dw_status st = prepare(m, recompute, s);
  if (st) return st;
  const int C = 3 * m->n + 2;
  const long chunk = m->chunk > 0 ? m->chunk : n_walkers;
  HIP_TRY(m->kappa_w.ensure((size_t)n_walkers * 4));
  m->kappa_n = n_walkers;
  // flags drive the recompute decision even when the caller does not ask for them
  uint32_t* flags = out_flags;
  if (recompute && !flags) {
    HIP_TRY(m->own_flags.ensure((size_t)n_walkers * 4));
    flags = m->own_flags.as<uint32_t>();
  }
  for (long b0 = 0; b0 < n_walkers; b0 += chunk) {
    const long nb = std::min<long>(chunk, n_walkers - b0);
    st = pass(m, r + b0 * m->n * 3, nb, R, Z, C, s, out_log_psi_sqr ? out_log_psi_sqr + b0 : nullptr, nullptr,
              out_eloc + b0, flags ? flags + b0 : nullptr, m->kappa_w.as<float>() + b0, m->d.fp64_cond_threshold);
    if (st) return st;
  }
  if (!recompute) return DW_OK;
  // FP64 recompute of the walkers whose conditioning exceeds the threshold (one host synchronisation)
  std::vector<uint32_t> hf(n_walkers);
  HIP_TRY(hipMemcpyAsync(hf.data(), flags, (size_t)n_walkers * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::vector<int32_t> idx;
  for (int b = 0; b < n_walkers; ++b)
    if (hf[b] & DW_FLAG_FP64_DET) idx.push_back(b);
  if (idx.empty()) return DW_OK;
  const int cntw = (int)idx.size();
  HIP_TRY(m->sub_idx.ensure((size_t)cntw * 4));
  HIP_TRY(m->sub_r.ensure((size_t)cntw * m->n * 3 * 4));
  HIP_TRY(m->sub_e.ensure((size_t)cntw * 4));
  HIP_TRY(m->sub_lp.ensure((size_t)cntw * 4));
  HIP_TRY(hipMemcpyAsync(m->sub_idx.p, idx.data(), (size_t)cntw * 4, hipMemcpyHostToDevice, s));
  launch_gather_walkers(r, m->sub_idx.as<int32_t>(), cntw, m->n, m->sub_r.as<float>(), s);
  const long c2 = m->chunk > 0 ? m->chunk : cntw;
  for (long b0 = 0; b0 < cntw; b0 += c2) {
    const long nb = std::min<long>(c2, cntw - b0);
    PassBufs<double> pb;
    st = pass_scratch(m, nb, C, &pb);
    if (st) return st;
    // the conditioning record keeps the FP32 estimate that flagged the walker
    st = run_pass<double>(m, pb, m->sub_r.as<float>() + b0 * m->n * 3, nb, R, Z, C, s, m->sub_lp.as<float>() + b0,
                          nullptr, m->sub_e.as<float>() + b0, nullptr, nullptr, 0.f);
    if (st) return st;
  }
  launch_scatter_values(m->sub_e.as<float>(), m->sub_idx.as<int32_t>(), cntw, out_eloc, s);
  if (out_log_psi_sqr) launch_scatter_values(m->sub_lp.as<float>(), m->sub_idx.as<int32_t>(), cntw, out_log_psi_sqr, s);
  // the host index list is copied before this call returns
  HIP_TRY(hipStreamSynchronize(s));
  return DW_OK;
}

dw_status dw_clip_energies(const float* e_loc, int32_t n_walkers, int32_t clip_tanh, int32_t center_median,
                           int32_t width_mae, float clip_by, int32_t from_previous_step, const double* state_in,
                           float* out_weights, float* out_e_clipped, double* out_stats, dw_stream stream) {
  if (!e_loc || !out_weights || !out_stats) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  if (center_median && n_walkers > clip_max_walkers())
    return fail(DW_ERR_INVALID, "a median clipping center supports at most " + std::to_string(clip_max_walkers()) +
                                    " walkers per device");
  launch_clip_energies(e_loc, n_walkers, clip_tanh ? 1 : 0, center_median ? 1 : 0, width_mae ? 1 : 0, (double)clip_by,
                       from_previous_step ? 1 : 0, state_in, out_weights, out_e_clipped, out_stats,
                       (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

int64_t dw_param_count(const dw_model* m) {
  if (!m) return 0;
  int64_t tot = 0;
  for (auto& e : m->expected) {
    int64_t cnt = 1;
    for (auto v : e.second) cnt *= v;
    tot += cnt;
  }
  return tot;
}

dw_status dw_param_layout(const dw_model* m, int32_t index, char* key, int32_t key_len, int64_t* offset,
                          int64_t* shape) {
  if (!m || !key || key_len <= 0 || !offset || !shape) return fail(DW_ERR_INVALID, "null argument");
  if (index < 0 || index >= (int32_t)m->expected.size()) return fail(DW_ERR_INVALID, "leaf index out of range");
  const auto& e = m->expected[index];
  std::snprintf(key, (size_t)key_len, "%s", e.first.c_str());
  *offset = leaf_offset(m, e.first);
  shape[0] = e.second.size() > 0 ? e.second[0] : 1;
  shape[1] = e.second.size() > 1 ? e.second[1] : 1;
  return DW_OK;
}

dw_status dw_param_grad(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                        const float* walker_weights, float* out_grad, float* out_log_psi_sqr, dw_stream stream) {
  if (!m || !r || !R || !Z || !walker_weights || !out_grad) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  if (m->n > 52) return fail(DW_ERR_INVALID, "dw_param_grad supports up to 52 electrons (the FP64 inverse in LDS)");
  hipStream_t s = (hipStream_t)stream;
  dw_status st = prepare(m, false, s);
  if (st) return st;
  if (m->prec == GEMM_F64)
    return run_grad<double>(m, r, n_walkers, R, Z, walker_weights, out_grad, out_log_psi_sqr, s);
  return run_grad<float>(m, r, n_walkers, R, Z, walker_weights, out_grad, out_log_psi_sqr, s);
}

dw_status dw_threefry_split(const uint32_t* key, int32_t num, uint32_t* out, dw_stream stream) {
  if (!key || !out) return fail(DW_ERR_INVALID, "null argument");
  if (num <= 0) return fail(DW_ERR_INVALID, "num must be positive");
  launch_threefry_split(key, num, out, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

dw_status dw_threefry_normal(const uint32_t* key, int64_t count, float* out, dw_stream stream) {
  if (!key || !out) return fail(DW_ERR_INVALID, "null argument");
  if (count <= 0 || count > 0xFFFFFFFFLL) return fail(DW_ERR_INVALID, "count must be in [1, 2^32)");
  launch_threefry_normal(key, (long)count, out, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

dw_status dw_threefry_mcmc_draws(const uint32_t* keys_in, uint32_t* keys_out, int32_t n_walkers, int32_t n_el,
                                 int32_t n_steps, float* normals, float* uniforms, dw_stream stream) {
  if (!keys_in || !keys_out || !normals || !uniforms) return fail(DW_ERR_INVALID, "null argument");
  if (keys_in == keys_out) return fail(DW_ERR_INVALID, "keys_out must not alias keys_in");
  if (n_walkers <= 0 || n_el <= 0 || n_steps <= 0) return fail(DW_ERR_INVALID, "bad sizes");
  launch_threefry_mcmc(keys_in, keys_out, n_walkers, n_el, n_steps, normals, uniforms, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

dw_status dw_model_set_proposal(dw_model* m, int32_t kind, float r_min, float r_max, float langevin_scale) {
  if (!m) return fail(DW_ERR_INVALID, "null model");
  if (kind < DW_PROPOSAL_NORMAL || kind > DW_PROPOSAL_LOCAL_ONE_EL)
    return fail(DW_ERR_INVALID, "unknown proposal kind");
  if ((kind == DW_PROPOSAL_LOCAL || kind == DW_PROPOSAL_LANGEVIN || kind == DW_PROPOSAL_LOCAL_ONE_EL) &&
      !(r_min > 0.f && r_max >= r_min && std::isfinite(r_max) && std::isfinite(langevin_scale)))
    return fail(DW_ERR_INVALID, "need 0 < r_min <= r_max (finite) and a finite langevin_scale");
  m->prop_kind = kind;
  m->prop_rmin = r_min;
  m->prop_rmax = r_max;
  m->prop_scale = langevin_scale;
  return DW_OK;
}

dw_status dw_mcmc_propose(dw_model* m, const float* r, int32_t n_walkers, const float* R, const float* Z,
                          const float* normals, const dw_chain_scalars* chain, float* out_r, float* out_log_q,
                          dw_stream stream) {
  if (!m || !r || !R || !Z || !normals || !chain || !out_r || !out_log_q) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  if (out_r == r) return fail(DW_ERR_INVALID, "out_r must not alias r");
  const hipStream_t s = (hipStream_t)stream;
  const int kind = m->prop_kind;
  if (kind == DW_PROPOSAL_NORMAL || kind == DW_PROPOSAL_NORMAL_ONE_EL || kind == DW_PROPOSAL_LOCAL_ONE_EL) {
    if (kind == DW_PROPOSAL_NORMAL)
      launch_propose(r, normals, chain, out_r, n_walkers * m->n * 3, s);
    if (kind != DW_PROPOSAL_LOCAL_ONE_EL)
      launch_fill_i32((int32_t*)out_log_q, 0, n_walkers, s);  // +0.0f; a kernel, so the call can be captured
    if (kind != DW_PROPOSAL_NORMAL)
      launch_propose_one_el(kind, r, normals, chain, R, m->I, m->prop_rmin, m->prop_rmax, out_r, out_log_q, n_walkers,
                            m->n, s);
  } else {
    launch_propose_scaled(m->prop_kind, r, normals, chain, R, Z, m->I, m->prop_rmin, m->prop_rmax, m->prop_scale,
                          out_r, out_log_q, n_walkers, m->n, s);
  }
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

dw_status dw_energy_moments(const float* e_loc, int32_t n_walkers, double* out_mean_var, dw_stream stream) {
  if (!e_loc || !out_mean_var) return fail(DW_ERR_INVALID, "null argument");
  if (n_walkers <= 0) return fail(DW_ERR_INVALID, "n_walkers must be positive");
  launch_energy_moments(e_loc, n_walkers, out_mean_var, (hipStream_t)stream);
  HIP_TRY(hipGetLastError());
  return DW_OK;
}

dw_status dw_mcmc_steps(dw_model* m, float* r, int32_t n_walkers, const float* R, const float* Z, float* log_psi_sqr,
                        int32_t* walker_age, const float* normals, const float* uniforms, int32_t n_steps,
                        dw_chain_scalars* chain, int32_t max_age, int32_t stepsize_update_interval,
                        float target_acceptance, float min_stepsize, float max_stepsize, float* out_acc,
                        uint32_t* out_flags, dw_stream stream) {
  const McmcArgs a{r, n_walkers, R, Z, log_psi_sqr, walker_age, normals, uniforms, n_steps, chain, max_age,
                   stepsize_update_interval, target_acceptance, min_stepsize, max_stepsize, out_acc, out_flags};
  return dwmi::mcmc_run(m, a, nullptr, (hipStream_t)stream);
}

}  // extern "C"
