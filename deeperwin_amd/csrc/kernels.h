// Launchers of the dwmi device kernels (definitions in gemm.hip and wf_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dwmi {

// Epilogue of the batched-row GEMM  C[M, N] = A[M, K] * Bt[N, K]^T.
struct GemmEpi {
  const float* bias = nullptr;  // + bias[col] on rows with row % C == 0 (value channel)
  const float* V = nullptr;     // + V[((row / nC) * C + row % C) * ldv + col]   (per-walker broadcast)
  int ldv = 0;
  int C = 1;                    // channels per electron
  int nC = 1;                   // n_el * C
  int act_tanh = 0;             // y = tanh(.)  (value-only forward)
  int permute = 0;              // output rows go to spin-major order (up block, then down block)
  int n = 0, u = 0, B = 0;      // shape for `permute`
};

enum GemmPrec { GEMM_F32 = 0, GEMM_BF16X3 = 1, GEMM_BF16 = 2 };

// Operands: A row-major [M, lda] fp32, Bt row-major [>= pad128(N), ldb] fp32 (rows beyond N are zero),
// K a multiple of 32, lda/ldb multiples of 4.  For the BF16 forms, `Bhi`/`Blo` are the pre-split BF16
// weights with the same layout as Bt.
void gemm(GemmPrec prec, const float* A, int lda, const float* Bt, const uint16_t* Bhi, const uint16_t* Blo,
          int ldb, float* C, int ldc, int M, int N, int K, const GemmEpi& ep, hipStream_t s);

// B walkers, n electrons (u up), I ions, D dets, C channels per electron (1, or 3 n + 2 in the
// Laplacian pass) stored in Cp rows (C, or pad16(C) in the Laplacian pass), widths W1 / W2, L iterations.
struct Shape {
  int B, n, u, I, D, C, Cp, W1, W2, L;
};

// Fused 1-el GEMM + tanh forward-Laplacian rule of the Laplacian pass (rows: Cp per electron).
void gemm_lap(GemmPrec prec, const float* X, int ldx, const float* Bt, const uint16_t* Bhi, const uint16_t* Blo, int K,
              const float* V, int ldv, float* out, int ldo, int permute, const Shape& sh, hipStream_t s);

void launch_features0(const float* r, const float* R, float* X0, int Kx0, const Shape& sh, hipStream_t s);
void launch_twoel(const float* r, float* const* X, const int* Kx, const int* Kh, const float* W2e, const float* b2e,
                  const Shape& sh, hipStream_t s);
void launch_spin_mean(const float* X, int Kx, int Kh, float* MU, int Km, const Shape& sh, hipStream_t s);
void launch_envelope(const float* O, int ldo, const float* r, const float* R, const float* kappa,
                     const float* wenv, float* A, const Shape& sh, hipStream_t s);
void launch_det(const float* A, float* ld, float* sd, float* g, float* lapd, uint32_t* dflags, float cond_thr,
                const Shape& sh, hipStream_t s);
void launch_combine(const float* ld, const float* sd, const float* g, const float* lapd, const uint32_t* dflags,
                    const float* r, const float* R, const float* Z, float* out_lp, float* out_phase, float* out_eloc,
                    uint32_t* out_flags, const Shape& sh, hipStream_t s);

void launch_fill_i32(int32_t* p, int32_t v, int count, hipStream_t s);
void launch_propose(const float* r, const float* normals, const void* chain, float* r_prop, int count, hipStream_t s);
void launch_accept(float* r, const float* r_prop, float* lp, const float* lp_prop, const float* uniforms,
                   int32_t* age, int max_age, int32_t* acc_count, uint32_t* flags, int B, int n, hipStream_t s);
void launch_chain_update(void* chain, const int32_t* acc_count, int B, float* out_acc, int interval, float target,
                         float lo, float hi, hipStream_t s);

}  // namespace dwmi
