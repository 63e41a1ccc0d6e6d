// Launchers of the dwmi device kernels (definitions in gemm.hip and wf_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dwmi {

// Epilogue of the batched-row GEMM  C[M, N] = A[M, K] * Bt[N, K]^T.
template <typename T>
struct GemmEpiT {
  const T* bias = nullptr;      // + bias[col] on rows with row % C == 0 (value channel)
  const T* V = nullptr;         // + V[((row / nC) * C + row % C) * ldv + col]   (per-walker broadcast)
  int ldv = 0;
  int C = 1;                    // channels per electron
  int nC = 1;                   // n_el * C
  int act_tanh = 0;             // y = tanh(.)  (value-only forward)
  int permute = 0;              // output rows go to spin-major order (up block, then down block)
  int n = 0, u = 0, B = 0;      // shape for `permute`
  int vsplit = 1;               // V is the sum of vsplit partial buffers V + s * vstride
  long vstride = 0;
  int tile = -1;                // FP32 tile: -1 = heuristic, 0 = 64x64, 1 = 64x128, 2 = 128x128,
                                // 4 = 112x256 (N = 256, no split-K; else 64x64)
  int ksplit = 1;               // split K over ksplit partial outputs C + s * cstride (bias on s = 0 only)
  long cstride = 0;
};
using GemmEpi = GemmEpiT<float>;
using GemmEpiD = GemmEpiT<double>;

enum GemmPrec { GEMM_F32 = 0, GEMM_BF16X3 = 1, GEMM_BF16 = 2, GEMM_F64 = 3 };

// Operands: A row-major [M, lda] fp32, Bt row-major [>= pad128(N), ldb] fp32 (rows beyond N are zero),
// K a multiple of 32, lda/ldb multiples of 4.  For the BF16 forms, `Bhi`/`Blo` are the pre-split BF16
// weights with the same layout as Bt.
void gemm(GemmPrec prec, const float* A, int lda, const float* Bt, const uint16_t* Bhi, const uint16_t* Blo,
          int ldb, float* C, int ldc, int M, int N, int K, const GemmEpi& ep, hipStream_t s);

// FP64 form (VALU FMA, 64 x 64 tiles).  Same operand conventions; K a multiple of 16.
void gemm_f64(const double* A, int lda, const double* Bt, int ldb, double* C, int ldc, int M, int N, int K,
              const GemmEpiD& ep, hipStream_t s);

// B walkers, n electrons (u up), I ions, D dets, C channels per electron (1, or 3 n + 2 in the
// Laplacian pass) stored in Cp rows (C, or pad16(C) in the Laplacian pass), widths W1 / W2, L iterations.
struct Shape {
  int B, n, u, I, D, C, Cp, W1, W2, L;
};

// Fused 1-el GEMM + tanh forward-Laplacian rule of the Laplacian pass (rows: Cp per electron).
// E_override > 0 sets the electrons per workgroup (E * Cp / 16 <= 12); 0 picks lap_electrons_per_tile:
// FP32 min(4, 192 / Cp), BF16 forms 96 / Cp (at least 1).
// cols = 256 asks for 256-column tiles (FP32 form, E * Cp <= 96); otherwise 128.
void gemm_lap(GemmPrec prec, const float* X, int ldx, const float* Bt, const uint16_t* Bhi, const uint16_t* Blo, int K,
              const float* V, int ldv, float* out, int ldo, int permute, const Shape& sh, int E_override, int cols,
              hipStream_t s);
int lap_electrons_per_tile(int Cp, GemmPrec prec);

// Value-pass 1-el GEMM on the Laplacian GEMM's 16x16 MFMA main loop (FP32): out = tanh(X W + V[walker]),
// V summed over vsplit parts vstride apart; permute sends rows to spin-major order.
void gemm_value16(const float* X, int ldx, const float* Bt, int K, const float* V, int ldv, int vsplit, long vstride,
                  float* out, int ldo, int permute, const Shape& sh, hipStream_t s);

// Laplacian-pass per-walker GEMM with the spin means of X's h columns formed in LDS (FP32):
// V[r, o] = sum_k MU[r, k] Wmu[o, k] + bias[o] on value-channel rows, equal to launch_spin_mean followed by
// the bias-only gemm on MU.  Returns false, launching nothing, unless W1 == 256, Km >= 256, Km % 16 == 0 and
// Kh, Kx are multiples of 4 (C > 1).
bool gemm_mean_lap(const float* X, int Kx, int Kh, const float* Wmu, int Km, const float* bias, float* V, int ldv,
                   const Shape& sh, hipStream_t s);

// Layer-0 input rows: the el-ion features in columns [0, 4 I); the columns [tail0, tail0 + tailw) are left
// to the 2-el kernels for the real channel rows; every other column and padding row is zero.  With normals
// (and the device chain record), the rows are those of the Gaussian proposal r + stepsize * normals, which is
// also written to r_out.
template <typename T>
void launch_features0(const float* r, const float* R, T* X0, int Kx0, int tail0, int tailw, const Shape& sh,
                      hipStream_t s, const float* normals = nullptr, const void* chain = nullptr,
                      float* r_out = nullptr);
template <typename T>
void launch_twoel(const float* r, T* const* X, const int* Kx, const int* Kh, const T* W2e, const T* b2e,
                  const Shape& sh, hipStream_t s);
template <typename T>
void launch_spin_mean(const T* X, int Kx, int Kh, T* MU, int Km, const Shape& sh, hipStream_t s);
template <typename T>
void launch_tanh_rule(const T* Z, T* out, int ldo, int permute, const Shape& sh, hipStream_t s);
template <typename T>
void launch_envelope(const T* O, int ldo, const float* r, const float* R, const T* kappa, const T* wenv, T* A,
                     const Shape& sh, hipStream_t s);
// Orbitals and envelope parameters as the determinant kernels read them when they apply the envelope
// themselves (det_reads_orbitals): O holds spin-major rows (Cp per electron, ldo columns).
template <typename T>
struct OrbSrc {
  const T* O;
  int ldo;
  const float* r;
  const float* R;
  const T* kappa;
  const T* wenv;
};
// True when the determinant kernels for n electrons and C channels apply the envelope while reading O
// (no envelope launch, no stored product); otherwise launch_envelope fills A first.
inline bool det_reads_orbitals(int n, int C) { return C == 1 ? n <= 32 : n <= 16; }
// Outputs of the value-pass combine (launch_combine's, C == 1) when launch_det runs it itself.
struct CombineValue {
  float* out_lp;
  float* out_phase;
  uint32_t* out_flags;
  float* out_kappa;
  float cond_thr;
  int fuse;  // set by launch_det
};
// Returns true when it also ran the value-pass combine for `combine` (value pass, n <= 32, D dividing the
// determinants per workgroup); the caller then skips launch_combine.
template <typename TA>
bool launch_det(const TA* A, const OrbSrc<TA>& src, double* ld, double* sd, double* g, double* lapd, float* dcond,
                const Shape& sh, hipStream_t s, const CombineValue* combine = nullptr);
void launch_combine(const double* ld, const double* sd, const double* g, const double* lapd, const float* dcond,
                    float cond_thr, const float* r, const float* R, const float* Z, float* out_lp, float* out_phase,
                    float* out_eloc, uint32_t* out_flags, float* out_kappa, const Shape& sh, hipStream_t s);

void launch_fill_i32(int32_t* p, int32_t v, int count, hipStream_t s);
// JAX threefry streams (threefry.hip)
void launch_threefry_split(const uint32_t* key, int num, uint32_t* out, hipStream_t s);
void launch_threefry_normal(const uint32_t* key, long count, float* out, hipStream_t s);
void launch_threefry_mcmc(const uint32_t* keys_in, uint32_t* keys_out, int B, int n, int n_steps, float* normals,
                          float* uniforms, hipStream_t s);
// out[0] = nanmean(e[0:n]), out[1] = nanmean(|e - out[0]|^2), FP64, deterministic order.
void launch_energy_moments(const float* e, int n, double* out, hipStream_t s);
// Clipped-energy statistics and per-walker gradient weights (optimization/loss_function.py:12-110); one
// workgroup.  A median center needs n <= clip_max_walkers().
int clip_max_walkers();
void launch_clip_energies(const float* e, int n, int tanh_clip, int median, int mae, double clip_by,
                          int from_previous, const double* state_in, float* weights, float* eclip, double* stats,
                          hipStream_t s);
// ---- dpe4 embedding (csrc/dpe4.hip) ----
constexpr int MAXL_DPE4 = 4;  // iterations
// Pair streams: wm[l][b][i][j][ch][f] (ch: value, d/dd (3), total Laplacian in the Laplacian pass).
// Wp [2L-1][2][32][32], bp [2L-1][2][32]: the maps of iterations 0..L-1, then the stream layers.
template <typename T>
void launch_dpe4_pairs(const float* r, const T* Wp, const T* bp, T* wm, const Shape& sh, hipStream_t s);
// El-ion stream: e_ei of every iteration into X[l] columns tail[l] + 32 ...; him: scratch [L][I][32].
template <typename T>
void launch_dpe4_elion(const float* r, const float* R, const float* Z, const T* Wei, const T* bei, const T* Whim,
                       const T* bhim, const T* table, int n_types, int z_min, T* him, T* const* X, const int* Kx,
                       const int* tail, const Shape& sh, hipStream_t s);
// hm [B n Cp][32] = tanh rule of (hm + bias) over each electron's channel rows, in place.
template <typename T>
void launch_dpe4_tanh(T* hm, const T* bias, const Shape& sh, hipStream_t s);
// e_ee of one iteration from wm (that iteration's slice) and hm into X columns col0 .. col0 + 31.
template <typename T>
void launch_dpe4_conv(const T* wm, const T* hm, T* X, int Kx, int col0, const Shape& sh, hipStream_t s);

void launch_gather_walkers(const float* r, const int32_t* idx, int count, int n, float* out, hipStream_t s);
void launch_scatter_values(const float* v, const int32_t* idx, int count, float* out, hipStream_t s);
void launch_propose(const float* r, const float* normals, const void* chain, float* r_prop, int count, hipStream_t s);
// Local-stepsize (kind DW_PROPOSAL_LOCAL) and approximate-Langevin (DW_PROPOSAL_LANGEVIN) proposals
// (mcmc.py:218-236, 255-284): r_prop [B, n, 3] and the per-walker log proposal ratio logq [B].  n <= 64.
void launch_propose_scaled(int kind, const float* r, const float* normals, const void* chain, const float* R,
                           const float* Z, int I, float r_min, float r_max, float lscale, float* r_prop, float* logq,
                           int B, int n, hipStream_t s);
// One-electron moves (DW_PROPOSAL_NORMAL_ONE_EL / DW_PROPOSAL_LOCAL_ONE_EL, mcmc.py:183-192, 239-252): electron
// step_nr mod n moves by xi [B, 3]; logq [B] written by the local form only.
void launch_propose_one_el(int kind, const float* r, const float* xi, const void* chain, const float* R, int I,
                           float r_min, float r_max, float* r_prop, float* logq, int B, int n, hipStream_t s);
// logq (may be null for a symmetric proposal) is added to the exponent: accept iff exp(lp' - lp + logq) > u.
// The step's accepted walkers are added to *acc_count; the last workgroup (by *ticket, which must be zero
// on entry) then updates the chain record: step_nr, acc_rate and, every `interval` steps, the stepsize from
// the previous acc_rate, clipped to [lo, hi]; out_acc (may be null) receives the step's acceptance.
void launch_accept(float* r, const float* r_prop, float* lp, const float* lp_prop, const float* logq,
                   const float* uniforms, int32_t* age, int max_age, int32_t* acc_count, uint32_t* flags, int B,
                   int n, void* chain, int32_t* ticket, float* out_acc, int interval, float target, float lo,
                   float hi, hipStream_t s);

}  // namespace dwmi
