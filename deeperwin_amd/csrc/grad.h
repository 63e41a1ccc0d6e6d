// Launchers of the parameter-gradient (reverse-mode) kernels in grad.hip.  They run after a value pass
// whose activations are kept, and accumulate G = sum_b c_b d log|psi(r_b)|^2 / d theta.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace dwmi {

// Seed of the backward pass: gl[b, d] = c_b d log_psi_sqr / d l_d, differentiated literally through the max
// shift and the epsilon inside the log (model/wavefunction.py:63-83); l_d, s_d from the value pass.
void launch_grad_seed(const double* ld, const double* sd, const float* c, double* gl, const Shape& sh, hipStream_t s);

// Per (walker, det): the orbital matrix A_d (envelope x backflow, as the value pass forms it), its inverse in
// FP64, and dL/dA_d = gl A_d^{-T}; written as dL/d backflow (times the envelope) and dL/d envelope (times the
// backflow), in the spin-major row order of O, ld columns.
template <typename T>
void launch_det_bwd(const OrbSrc<T>& src, const double* gl, T* dbf, T* denv, int ld, const Shape& sh, hipStream_t s);

// Envelope parameter gradients: for every (spin, ion, column), partial sums over walker chunks of
//   dalpha = sum denv w e^{-kappa rho} (-rho) (1 - e^{-kappa}),  dw = sum denv e^{-kappa rho}.
// part: [chunks, 2 (alpha, w), 2 (spin), I, NO].  Returns the number of chunks.
template <typename T>
int launch_env_grad(const T* denv, int ld, const float* r, const float* R, const T* kappa, const T* wenv, T* part,
                    int walkers_per_chunk, const Shape& sh, hipStream_t s);

// dz of the last layer: dz[b n + i, o] = dH[orb_row(b, i), o] (1 - H[orb_row, o]^2)   (electron order out)
template <typename T>
void launch_dz_last(const T* dH, const T* H, T* dz, const Shape& sh, hipStream_t s);
// dz of layer l - 1 from the gradient of its output h^l (columns 0..Kh-1 of X_l):
//   dh = dX[:, o] + dMU[b, s(i) Kh + o] / n_s(i),   dz = dh (1 - h^2)
template <typename T>
void launch_dz(const T* dX, int ldx, const T* dMU, int ldmu, const T* X, int ldX, int Kh, T* dz, const Shape& sh,
               hipStream_t s);
// Per-walker sums over electrons: out[b, o] = sum_i in[b n + i, o]
template <typename T>
void launch_walker_sum(const T* in, int ld, int ncols, T* out, const Shape& sh, hipStream_t s);
// Column sums in fixed order: out[o] (float) = sum_r in[r, o]; part holds >= colsum_chunks(rows) * ncols values.
int colsum_chunks(long rows);
template <typename T>
void launch_colsum(const T* in, int ld, long rows, int ncols, T* part, float* out, hipStream_t s);

template <typename T>
struct TwoElBwdArgs {
  const T* dX[4];  // dX_l (gradient of the input of 1-el layer l), row pitch ldx[l], pair means from column Kh[l]
  int ldx[4];
  int Kh[4];
  T* X2[3];  // inputs of 2-el layer l, [pairs, 4 or 32]
  T* DP[3];  // pre-activation gradients of 2-el layer l, [pairs, 32]
  const T* W2e;  // [L - 1][32][32] (input x output, rows beyond the input width zero)
  const T* b2e;  // [L - 1][32]
};

// 2-el stream backward: per pair (b, i, j), recompute the 2-el MLP and push the pair-mean gradients
// dX_l[:, Kh_l .. Kh_l + 2 Ke_l) (l >= 1) back through it.  Writes the inputs X2_l [pairs, kin_l] and the
// pre-activation gradients DP_l [pairs, 32] of 2-el layers l = 0 .. L-2.
template <typename T>
void launch_twoel_bwd(const float* r, T* const* dX, const int* Kx, const int* Kh, const T* W2e, const T* b2e,
                      T* const* X2, T* const* DP, const Shape& sh, hipStream_t s);

// ---- dpe4 embedding backward (dpe4_grad.hip) ----
// dhm[b n + j][f] = sum_i de[b n + i][f] wm[b][i][j][f], dw[b][i][j][f] = de[b n + i][f] hm[b n + j][f], with de
// the columns col0 .. col0 + 31 of dX (row pitch ldx).
template <typename T>
void launch_dpe4_conv_bwd(const T* dX, int ldx, int col0, const T* wm, const T* hm, T* dhm, T* dw, const Shape& sh,
                          hipStream_t s);
// d[t] *= 1 - y[t]^2
template <typename T>
void launch_dtanh(T* d, const T* y, long count, hipStream_t s);
// dst[r, c] += src[r, c] for c < ncols
template <typename T>
void launch_add_cols(T* dst, int ldd, const T* src, int lds, long rows, int ncols, hipStream_t s);
// Pair streams backward: per pair, from dw_l (natural [b][i][j][32] order) and the maps wm_l of the value pass,
// the stream inputs X2[l], map pre-activation gradients DZm[l] and stream-layer gradients DZh[l] (l < L-1), in
// rows grouped by spin relation (same-spin pairs of every walker, then the diff-spin pairs).
template <typename T>
void launch_dpe4_pairs_bwd(const float* r, const T* Wp, const T* bp, T* const* dw, T* const* wm, T* const* X2,
                           T* const* DZm, T* const* DZh, const Shape& sh, hipStream_t s);
// El-ion stream backward per (electron row, ion): de_ei_l from dX[l] columns col[l] ..; writes the stream inputs
// Q[l], stream-layer gradients DZQ[l] (l < L-1) and PH[l] = de_ei_l * q_l (to be summed over rows into dhim).
template <typename T>
void launch_dpe4_elion_bwd(const float* r, const float* R, const T* Wei, const T* bei, const T* him,
                           T* const* dX, const int* ldx, const int* col, T* const* Q, T* const* DZQ, T* const* PH,
                           const Shape& sh, hipStream_t s);
// Ion-map and ion-table gradients from dhim [L][I][32] (float); gW[l] / gB[l] / gTable are the leaves in the
// packed gradient.  One workgroup.
template <typename T>
void launch_dpe4_him_bwd(const float* dhim, const T* him, const T* table, const T* Whim, const float* Z, int z_min,
                         int n_types, int L, int I, float* const* gW, float* const* gB, float* gTable,
                         hipStream_t s);

// out[t] = (float) sum_c part[c * count + t], chunks in order.
template <typename T>
void launch_sum_parts(const T* part, int chunks, long count, float* out, hipStream_t s);

// Weight gradient C[m, n] = scale * sum_k A[k, m] B[k, n] over K rows (both operands row-major with the
// reduction index as the row), written as float to C with row pitch ldc.  part holds
// ceil(K / kchunk) * M * N values of T.
template <typename T>
void wgrad(const T* A, int lda, const T* Bm, int ldb, int M, int N, long K, T* part, float* C, int ldc,
           hipStream_t s);
size_t wgrad_part_elems(int M, int N, long K);

}  // namespace dwmi
