// Batched-row GEMMs on gfx950 MFMA with fused epilogues.
//
//   C[M, N] = epi(A[M, K] * Bt[N, K]^T)
//
// A holds one row per (walker, electron, channel); Bt is a layer weight stored output-major (K
// contiguous), so both operands stream K-contiguous vectors.  Tiles are up to 128 x 128 per 256-thread
// workgroup (4 waves in 2 x 2).  Tiles are assigned so that the N-tiles of one M-tile run back to back on
// the same XCD (shared L2 for the A rows).
//
// FP32 form: v_mfma_f32_16x16x4_f32.  BF16 forms: v_mfma_f32_32x32x16_bf16 with FP32 accumulation;
// A is split into hi + lo BF16 parts while it is staged into LDS, the weights are pre-split, and the
// x3 form accumulates hi*hi + hi*lo + lo*hi.
#include "common.h"
#include "kernels.h"

namespace dwmi {

enum { EPI_BIAS = 1, EPI_V = 2, EPI_TANH = 4, EPI_PERM = 8 };

__device__ __forceinline__ float act_tanh(float v) { return tanh_fast(v); }
__device__ __forceinline__ double act_tanh(double v) { return tanh(v); }

// Epilogue of one element at row r = (b * n + i) * C + ch, column c.  Index math in 32 bits (M < 2^31).
template <int FLAGS, typename T>
__device__ __forceinline__ void epi_store(T v, int r, int c, T* __restrict__ Cm, int ldc, const GemmEpiT<T>& ep) {
  int bi = r, ch = 0;
  if ((FLAGS & (EPI_BIAS | EPI_V | EPI_PERM)) && ep.C > 1) {
    bi = r / ep.C;
    ch = r - bi * ep.C;
  }
  if ((FLAGS & EPI_BIAS) && ep.bias) {
    if (ch == 0) v += ep.bias[c];
  }
  int b = 0;
  if ((FLAGS & (EPI_V | EPI_PERM)) && (ep.V || ep.permute)) b = bi / ep.n;
  if ((FLAGS & EPI_V) && ep.V) {
    const T* Vp = ep.V + ((long)b * ep.C + ch) * ep.ldv + c;
    T vs = Vp[0];
    for (int q = 1; q < ep.vsplit; ++q) vs += Vp[q * ep.vstride];
    v += vs;
  }
  if ((FLAGS & EPI_TANH) && ep.act_tanh) v = act_tanh(v);
  long orow = r;
  if ((FLAGS & EPI_PERM) && ep.permute) {
    const int i = bi - b * ep.n, dn = ep.n - ep.u;
    orow = (i < ep.u) ? ((long)b * ep.u + i) * ep.C + ch
                      : (long)ep.B * ep.u * ep.C + ((long)b * dn + (i - ep.u)) * ep.C + ch;
  }
  Cm[orow * ldc + c] = v;
}

__device__ __forceinline__ int tile_index(int bid, int total) {
  // XCD-aware: workgroups are dealt round-robin to 8 XCDs; give each XCD a contiguous tile range.
  const int per = (total + 7) / 8;
  return (bid & 7) * per + (bid >> 3);
}

// ------------------------------------------------------------------------------------------------
// FP32 MFMA.  Tile (64 WM) x (64 WN) per 256-thread workgroup; 4 waves in 2 x 2, each wave owns
// (2 WM) x (2 WN) blocks of 16 x 16 (v_mfma_f32_16x16x4_f32), so even the 64 x 64 tile keeps four
// independent accumulators per wave (32 x 32 blocks would leave it one dependent chain).  WM = WN = 2 for
// large GEMMs; smaller tiles give enough workgroups for the per-walker and orbital GEMMs of a value pass.
// Every tile shape adds the same products to an output element in the same order, so the tile choice
// never changes a result.  Both operand slices (16 k) are staged
// untransposed in LDS with a row pitch of 20 floats; MFMA step s of lane group g = lane / 16 takes
// k = 4 g + s, so a lane reads its operands for a whole slice as one 16-byte read per block (the 16 rows
// one read touches start 20 floats apart and cover disjoint banks) and waits on LDS once per slice.
// ------------------------------------------------------------------------------------------------
template <int FLAGS, int WM, int WN, int BK>
__global__ __launch_bounds__(256) void k_gemm_f32(const float* __restrict__ A, int lda, const float* __restrict__ Bt,
                                                  int ldb, float* __restrict__ Cm, int ldc, int M, int N, int K,
                                                  int tilesN, int tilesTotal, GemmEpi ep) {
  constexpr int BM = 64 * WM, BN = 64 * WN, PITCH = BK + 4;  // BK = 16 or 32
  constexpr int KV = BK / 4;  // float4 per row of a K-slice
  constexpr int AV = BM * KV, BV = BN * KV;  // float4 per tile
  constexpr int AIT = (AV + 255) / 256, BIT = (BV + 255) / 256;
  constexpr int MB = 2 * WM, NB = 2 * WN;  // 16 x 16 blocks per wave
  __shared__ __align__(16) float As[BM][PITCH];
  __shared__ __align__(16) float Bs[BN][PITCH];
  // split-K: workgroup range [split * gridTiles, (split + 1) * gridTiles) computes K slice `split`
  const int gridTiles = (tilesTotal + 7) / 8 * 8;
  const int split = blockIdx.x / gridTiles;
  const int tile = tile_index(blockIdx.x - split * gridTiles, tilesTotal);
  if (tile >= tilesTotal) return;
  const int Ks = K / ep.ksplit;
  A += (long)split * Ks;
  Bt += (long)split * Ks;
  Cm += split * ep.cstride;
  K = Ks;
  const long row0 = (long)(tile / tilesN) * BM;
  const int col0 = (tile % tilesN) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid >> 1, wn = wid & 1;
  const int g = lane >> 4, cl = lane & 15;

  v4f acc[MB][NB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};

  // staging: vector index idx -> row 16 (idx / (16 KV)) + idx % 16, k quad (idx / 16) % KV, so the 16
  // lanes of one store phase write 16 rows at one quad (the banks of the 16-byte reads: no conflicts)
  auto srow = [](int idx) { return (idx / (16 * KV)) * 16 + (idx & 15); };
  auto squad = [](int idx) { return ((idx >> 4) % KV) * 4; };
  v4f ra[AIT], rb[BIT];
  auto gload = [&](int k0) {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      const int idx = tid + 256 * it;
      if (idx < AV) {
        long gr = row0 + srow(idx);
        if (gr > M - 1) gr = M - 1;
        ra[it] = *(const v4f*)(A + gr * lda + k0 + squad(idx));
      }
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      const int idx = tid + 256 * it;
      if (idx < BV) rb[it] = *(const v4f*)(Bt + (long)(col0 + srow(idx)) * ldb + k0 + squad(idx));
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      const int idx = tid + 256 * it;
      if (idx < AV) *(v4f*)&As[srow(idx)][squad(idx)] = ra[it];
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      const int idx = tid + 256 * it;
      if (idx < BV) *(v4f*)&Bs[srow(idx)][squad(idx)] = rb[it];
    }
  };

  gload(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {  // 16-wide sub-slices in order: the same sequence for any BK
      v4f a4[MB], b4[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) a4[i] = *(const v4f*)&As[wm * 32 * WM + i * 16 + cl][kk + 4 * g];
#pragma unroll
      for (int j = 0; j < NB; ++j) b4[j] = *(const v4f*)&Bs[wn * 32 * WN + j * 16 + cl][kk + 4 * g];
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[i][st], b4[j][st], acc[i][j], 0, 0, 0);
    }
  }

#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = (int)row0 + wm * 32 * WM + i * 16 + 4 * g + e;
        const int c = col0 + wn * 32 * WN + j * 16 + cl;
        if (r < M && c < N) {
          if (split == 0)
            epi_store<FLAGS, float>(acc[i][j][e], r, c, Cm, ldc, ep);
          else
            Cm[(long)r * ldc + c] = acc[i][j][e];
        }
      }
}

// ------------------------------------------------------------------------------------------------
// BF16 MFMA (x1 and x3 split forms)
// ------------------------------------------------------------------------------------------------
template <int FLAGS, int X3>
__global__ __launch_bounds__(256) void k_gemm_bf16(const float* __restrict__ A, int lda,
                                                   const uint16_t* __restrict__ Bhi, const uint16_t* __restrict__ Blo,
                                                   int ldb, float* __restrict__ Cm, int ldc, int M, int N, int K,
                                                   int tilesN, int tilesTotal, GemmEpi ep) {
  constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;  // row pitch 80 B
  __shared__ uint16_t Ah[BM][LDK];
  __shared__ uint16_t Al[X3 ? BM : 1][LDK];
  __shared__ uint16_t Bh[BN][LDK];
  __shared__ uint16_t Bl[X3 ? BN : 1][LDK];
  // split-K: workgroup range [split * gridTiles, (split + 1) * gridTiles) computes K slice `split`
  const int gridTiles = (tilesTotal + 7) / 8 * 8;
  const int split = blockIdx.x / gridTiles;
  const int tile = tile_index(blockIdx.x - split * gridTiles, tilesTotal);
  if (tile >= tilesTotal) return;
  const int Ks = K / ep.ksplit;
  A += (long)split * Ks;
  Bhi += (long)split * Ks;  // (the BF16 launcher never splits K: split == 0)
  if (Blo) Blo += (long)split * Ks;
  Cm += split * ep.cstride;
  K = Ks;
  const long row0 = (long)(tile / tilesN) * BM;
  const int col0 = (tile % tilesN) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid >> 1, wn = wid & 1;

  v16f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[a][b][j] = 0.f;

  // A: 128 x 32 fp32 = 1024 float4 -> 4 per thread.  B: 128 x 32 bf16 = 1024 x (8 B) -> 4 per thread.
  float4 ra[4];
  uint2 rbh[4], rbl[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + 256 * it, m = idx >> 3, kq = (idx & 7) * 4;
      long gr = row0 + m;
      if (gr > M - 1) gr = M - 1;
      ra[it] = *(const float4*)(A + gr * lda + k0 + kq);
    }
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + 256 * it, m = idx >> 3, kq = (idx & 7) * 4;
      rbh[it] = *(const uint2*)(Bhi + (long)(col0 + m) * ldb + k0 + kq);
      if (X3) rbl[it] = *(const uint2*)(Blo + (long)(col0 + m) * ldb + k0 + kq);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + 256 * it, m = idx >> 3, kq = (idx & 7) * 4;
      const float f[4] = {ra[it].x, ra[it].y, ra[it].z, ra[it].w};
      unsigned short h[4], l[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        h[q] = f2bf_bits(f[q]);
        l[q] = f2bf_bits(f[q] - bf_bits2f(h[q]));
      }
      uint2 hv, lv;
      hv.x = h[0] | ((unsigned)h[1] << 16);
      hv.y = h[2] | ((unsigned)h[3] << 16);
      lv.x = l[0] | ((unsigned)l[1] << 16);
      lv.y = l[2] | ((unsigned)l[3] << 16);
      *(uint2*)&Ah[m][kq] = hv;
      if (X3) *(uint2*)&Al[m][kq] = lv;
    }
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int idx = tid + 256 * it, m = idx >> 3, kq = (idx & 7) * 4;
      *(uint2*)&Bh[m][kq] = rbh[it];
      if (X3) *(uint2*)&Bl[m][kq] = rbl[it];
    }
  };

  gload(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kc = kk + 8 * (lane >> 5);
      v8s ah[2], bh[2], al[2], bl[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        ah[q] = *(const v8s*)&Ah[wm * 64 + q * 32 + (lane & 31)][kc];
        bh[q] = *(const v8s*)&Bh[wn * 64 + q * 32 + (lane & 31)][kc];
        if (X3) {
          al[q] = *(const v8s*)&Al[wm * 64 + q * 32 + (lane & 31)][kc];
          bl[q] = *(const v8s*)&Bl[wn * 64 + q * 32 + (lane & 31)][kc];
        }
      }
#pragma unroll
      for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
          if (X3) {
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16((v8bf)al[mi], (v8bf)bh[ni], acc[mi][ni], 0, 0, 0);
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16((v8bf)ah[mi], (v8bf)bl[ni], acc[mi][ni], 0, 0, 0);
          }
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16((v8bf)ah[mi], (v8bf)bh[ni], acc[mi][ni], 0, 0, 0);
        }
    }
  }

#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int r = (int)row0 + wm * 64 + mi * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        const int c = col0 + wn * 64 + ni * 32 + (lane & 31);
        if (r < M && c < N) epi_store<FLAGS, float>(acc[mi][ni][j], r, c, Cm, ldc, ep);
      }
}

// ------------------------------------------------------------------------------------------------
// FP64 (VALU FMA): 64 x 64 tile per 256-thread workgroup, thread (ty, tx) owns rows ty + 16 i and
// columns tx + 16 j (i, j < 4), BK = 16.
// ------------------------------------------------------------------------------------------------
template <int FLAGS>
__global__ __launch_bounds__(256) void k_gemm_f64(const double* __restrict__ A, int lda, const double* __restrict__ Bt,
                                                  int ldb, double* __restrict__ Cm, int ldc, int M, int N, int K,
                                                  int tilesN, int tilesTotal, GemmEpiD ep) {
  constexpr int BM = 64, BN = 64, BK = 16;
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  const int tile = tile_index(blockIdx.x, tilesTotal);
  if (tile >= tilesTotal) return;
  const long row0 = (long)(tile / tilesN) * BM;
  const int col0 = (tile % tilesN) * BN;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int idx = tid + 256 * it, m = idx >> 3, kq = (idx & 7) * 2;
      long gr = row0 + m;
      if (gr > M - 1) gr = M - 1;
      const double2 a2 = *(const double2*)(A + gr * lda + k0 + kq);
      As[kq][m] = a2.x;
      As[kq + 1][m] = a2.y;
      long gc = col0 + m;
      const double2 b2 = *(const double2*)(Bt + gc * ldb + k0 + kq);
      Bs[kq][m] = b2.x;
      Bs[kq + 1][m] = b2.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = (int)row0 + ty + 16 * i;
      const int c = col0 + tx + 16 * j;
      if (r < M && c < N) epi_store<FLAGS, double>(acc[i][j], r, c, Cm, ldc, ep);
    }
}

void gemm_f64(const double* A, int lda, const double* Bt, int ldb, double* C, int ldc, int M, int N, int K,
              const GemmEpiD& ep, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  const int tilesM = (M + 63) / 64, tilesN = (N + 63) / 64;
  const int total = tilesM * tilesN;
  const int grid = (total + 7) / 8 * 8;
  // one generic instance: flags are tested at run time inside epi_store
  hipLaunchKernelGGL((k_gemm_f64<EPI_BIAS | EPI_V | EPI_TANH | EPI_PERM>), dim3(grid), dim3(256), 0, s, A, lda, Bt,
                     ldb, C, ldc, M, N, K, tilesN, total, ep);
}

template <int FLAGS, int WM, int WN>
static void launch_f32(const float* A, int lda, const float* Bt, int ldb, float* C, int ldc, int M, int N, int K,
                       const GemmEpi& ep, hipStream_t s) {
  const int tilesM = (M + 64 * WM - 1) / (64 * WM), tilesN = (N + 64 * WN - 1) / (64 * WN);
  const int total = tilesM * tilesN;
  const int grid = (total + 7) / 8 * 8 * ep.ksplit;
  // 32-wide K slices on 128 x 128 tiles (half the barriers per MFMA; orbital GEMMs of the N2 Laplacian pass
  // 940 -> 845 us); the small tiles measured faster with 16 (N2 value pass 448 vs 485 us).  Results are the
  // same either way: a slice runs as 16-wide sub-slices in order.
  if (WM == 2 && WN == 2 && (K / ep.ksplit) % 32 == 0)
    hipLaunchKernelGGL((k_gemm_f32<FLAGS, WM, WN, 32>), dim3(grid), dim3(256), 0, s, A, lda, Bt, ldb, C, ldc, M, N,
                       K, tilesN, total, ep);
  else
    hipLaunchKernelGGL((k_gemm_f32<FLAGS, WM, WN, 16>), dim3(grid), dim3(256), 0, s, A, lda, Bt, ldb, C, ldc, M, N,
                       K, tilesN, total, ep);
}

// ------------------------------------------------------------------------------------------------
// FP32 MFMA on full-width row tiles: 16 RB rows (112 for the value pass, 128 for the large orbital GEMMs)
// x all columns (N <= 256) per 1024-thread workgroup (the A tile is read once; at 28,672 rows and RB = 7,
// one workgroup per CU), wave w owning columns [16 w, 16 w + 16) of all rows (RB blocks of 16 x 16; four
// waves per SIMD hide each other's LDS and barrier waits: measured faster than 8 waves of 7 x 2 blocks and
// 4 waves of 7 x 4 blocks); waves whose columns lie past N only stage operands.  Both operand slices are
// double-buffered in LDS (one barrier per 16-wide slice) with the next slice's loads in registers.  The
// slices, the LDS layout and the MFMA step order (step s of lane group g takes k = 4 g + s) are
// k_gemm_f32's, and so is the epilogue: results are bit-identical to it.  128 < N <= 256, N % 16 == 0 (Bt
// then holds 256 rows), K % 16 == 0, no split-K; with a V epilogue, RB = 7, N = 256, C == 1 and n >= 8.
// ------------------------------------------------------------------------------------------------
template <int FLAGS, int RB>
__global__ __launch_bounds__(1024) void k_gemm_rows256(const float* __restrict__ A, int lda,
                                                      const float* __restrict__ Bt, int ldb, float* __restrict__ Cm,
                                                      int ldc, int M, int N, int K, int tilesTotal, GemmEpi ep) {
  constexpr int NT = 1024;  // 16 waves, 4 per SIMD
  constexpr int BM = 16 * RB, BN = 256, BK = 16, PITCH = BK + 4, KV = BK / 4, NQ = 1;
  constexpr int AV = BM * KV, BV = BN * KV, AIT = (AV + NT - 1) / NT, BIT = BV / NT;
  __shared__ __align__(16) float As[2][BM][PITCH];
  __shared__ __align__(16) float Bs[2][BN][PITCH];
  const int tile = tile_index(blockIdx.x, tilesTotal);
  if (tile >= tilesTotal) return;  // uniform over the workgroup, before any barrier
  const long row0 = (long)tile * BM;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int g = lane >> 4, cl = lane & 15;

  v4f acc[RB][NQ];
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < NQ; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};

  auto srow = [](int idx) { return (idx / (16 * KV)) * 16 + (idx & 15); };
  auto squad = [](int idx) { return ((idx >> 4) % KV) * 4; };
  v4f ra[AIT], rb[BIT];
  auto gload = [&](int k0) {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      const int idx = tid + NT * it;
      if (idx < AV) {
        long gr = row0 + srow(idx);
        if (gr > M - 1) gr = M - 1;
        ra[it] = *(const v4f*)(A + gr * lda + k0 + squad(idx));
      }
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      const int idx = tid + NT * it;
      rb[it] = *(const v4f*)(Bt + (long)srow(idx) * ldb + k0 + squad(idx));
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int it = 0; it < AIT; ++it) {
      const int idx = tid + NT * it;
      if (idx < AV) *(v4f*)&As[buf][srow(idx)][squad(idx)] = ra[it];
    }
#pragma unroll
    for (int it = 0; it < BIT; ++it) {
      const int idx = tid + NT * it;
      *(v4f*)&Bs[buf][srow(idx)][squad(idx)] = rb[it];
    }
  };

  const int S = K / BK;
  const bool active = wid * 16 < N;  // uniform over the wave
  gload(0);
  sstore(0);
  if (S > 1) gload(BK);
  __syncthreads();
  for (int sl = 0; sl < S; ++sl) {
    const int buf = sl & 1;
    if (active) {
      v4f a4[RB], b4[NQ];
#pragma unroll
      for (int i = 0; i < RB; ++i) a4[i] = *(const v4f*)&As[buf][i * 16 + cl][4 * g];
#pragma unroll
      for (int j = 0; j < NQ; ++j) b4[j] = *(const v4f*)&Bs[buf][wid * 16 * NQ + j * 16 + cl][4 * g];
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < NQ; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[i][st], b4[j][st], acc[i][j], 0, 0, 0);
    }
    // the other buffer was last read before the previous barrier; the next slice's compute reads it
    // after this one
    if (sl + 1 < S) {
      sstore(buf ^ 1);
      if (sl + 2 < S) gload((sl + 2) * BK);
    }
    __syncthreads();
  }

  // Epilogue.  With V (the launcher guarantees one channel per electron and n >= 8, so a tile spans at most
  // 15 walkers) the walkers' V sums (parts added in epi_store's order) are formed once in LDS after the
  // main loop's last barrier, and the 112 outputs of a lane read no V from memory.
  constexpr int WMAX = 16;
  static_assert((FLAGS & EPI_V) == 0 || WMAX * BN <= 2 * BM * PITCH, "V staging fits in the A buffers");
  float(*Vs)[BN] = (float(*)[BN]) & As[0][0][0];
  int b_first = 0;
  if constexpr ((FLAGS & EPI_V) != 0) {
    const long r_last = (row0 + BM < M ? row0 + BM : (long)M) - 1;
    b_first = (int)(row0 / ep.n);
    const int nbw = (int)(r_last / ep.n) - b_first + 1;
    for (int idx = tid; idx < nbw * BN; idx += NT) {
      const int w = idx / BN, c = idx - w * BN;
      const float* Vp = ep.V + (long)(b_first + w) * ep.ldv + c;
      float vs = Vp[0];
      for (int q = 1; q < ep.vsplit; ++q) vs += Vp[q * ep.vstride];
      Vs[w][c] = vs;
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < NQ; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = (int)row0 + i * 16 + 4 * g + e;
        const int c = wid * 16 * NQ + j * 16 + cl;
        if (r < M && c < N) {
          if constexpr ((FLAGS & EPI_V) != 0) {
            // epi_store with C == 1: bias on every row, then V, tanh, spin-major permutation
            float v = acc[i][j][e];
            if ((FLAGS & EPI_BIAS) && ep.bias) v += ep.bias[c];
            const int b = r / ep.n;
            v += Vs[b - b_first][c];
            if ((FLAGS & EPI_TANH) && ep.act_tanh) v = act_tanh(v);
            long orow = r;
            if ((FLAGS & EPI_PERM) && ep.permute) {
              const int el = r - b * ep.n, dn = ep.n - ep.u;
              orow = (el < ep.u) ? (long)b * ep.u + el : (long)ep.B * ep.u + ((long)b * dn + (el - ep.u));
            }
            Cm[orow * ldc + c] = v;
          } else {
            epi_store<FLAGS, float>(acc[i][j][e], r, c, Cm, ldc, ep);
          }
        }
      }
}

// Workgroups needed to fill the chip a few times over: the largest tile that still gives >= 4 per CU.
static int pick_tile(long M, int N) {
  const long wg128 = ((M + 127) / 128) * ((N + 127) / 128);
  if (wg128 >= 1024) return 2;
  const long wg64x128 = ((M + 63) / 64) * ((N + 127) / 128);
  if (wg64x128 >= 1024) return 1;
  return 0;
}

template <int FLAGS>
static void gemm_dispatch(GemmPrec prec, const float* A, int lda, const float* Bt, const uint16_t* Bhi,
                          const uint16_t* Blo, int ldb, float* C, int ldc, int M, int N, int K, const GemmEpi& ep,
                          hipStream_t s) {
  if (prec == GEMM_F32) {
    // Bt holds pad128(N) rows, so every tile shape reads inside it.  Tile 4 (112 x 256 full-width rows,
    // N = 256 only) is what the heuristic takes where it would pick 64 x 64 tiles and the row tiles fill
    // at least 192 CUs (one workgroup per CU).
    int tile = ep.tile >= 0 ? ep.tile : pick_tile((long)M * ep.ksplit, N);
    if (ep.tile < 0 && tile == 0 && (M + 111) / 112 >= 192) tile = 4;
    // Plain GEMMs with 128 < N <= 256 and at least 1024 row tiles of 128 (the Laplacian-pass orbital GEMMs:
    // 688,128 x 256 x 224 at N2, 888 -> 816 us each): full-width 128-row tiles, waves past N idle in the
    // MFMA phase.
    if constexpr (FLAGS == 0) {
      if (ep.tile < 0 && N > 128 && N <= 256 && N % 16 == 0 && ep.ksplit == 1 && K % 16 == 0 && K >= 16 &&
          (M + 127) / 128 >= 1024) {
        const int total = (M + 127) / 128;
        hipLaunchKernelGGL((k_gemm_rows256<FLAGS, 8>), dim3((total + 7) / 8 * 8), dim3(1024), 0, s, A, lda, Bt, ldb,
                           C, ldc, M, N, K, total, ep);
        return;
      }
    }
    if (tile == 4 && !(N == 256 && ep.ksplit == 1 && K % 16 == 0 && K >= 16 && (!ep.V || (ep.C == 1 && ep.n >= 8))))
      tile = 0;
    if (tile == 4) {
      // (not for the generic all-flags instance, whose epilogue would not stay in registers)
      if constexpr (FLAGS != (EPI_BIAS | EPI_V | EPI_TANH | EPI_PERM)) {
        const int total = (M + 111) / 112;
        hipLaunchKernelGGL((k_gemm_rows256<FLAGS, 7>), dim3((total + 7) / 8 * 8), dim3(1024), 0, s, A, lda, Bt, ldb,
                           C, ldc, M, N, K, total, ep);
        return;
      }
      tile = 0;
    }
    switch (tile) {
      case 2: launch_f32<FLAGS, 2, 2>(A, lda, Bt, ldb, C, ldc, M, N, K, ep, s); break;
      case 1: launch_f32<FLAGS, 1, 2>(A, lda, Bt, ldb, C, ldc, M, N, K, ep, s); break;
      default: launch_f32<FLAGS, 1, 1>(A, lda, Bt, ldb, C, ldc, M, N, K, ep, s); break;
    }
    return;
  }
  const int tilesM = (M + 127) / 128, tilesN = (N + 127) / 128;
  const int total = tilesM * tilesN;
  const int grid = (total + 7) / 8 * 8;
  if (prec == GEMM_BF16X3)
    hipLaunchKernelGGL((k_gemm_bf16<FLAGS, 1>), dim3(grid), dim3(256), 0, s, A, lda, Bhi, Blo, ldb, C, ldc, M, N, K,
                       tilesN, total, ep);
  else
    hipLaunchKernelGGL((k_gemm_bf16<FLAGS, 0>), dim3(grid), dim3(256), 0, s, A, lda, Bhi, Blo, ldb, C, ldc, M, N, K,
                       tilesN, total, ep);
}

void gemm(GemmPrec prec, const float* A, int lda, const float* Bt, const uint16_t* Bhi, const uint16_t* Blo, int ldb,
          float* C, int ldc, int M, int N, int K, const GemmEpi& ep, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  const int flags = (ep.bias ? EPI_BIAS : 0) | (ep.V ? EPI_V : 0) | (ep.act_tanh ? EPI_TANH : 0) |
                    (ep.permute ? EPI_PERM : 0);
  switch (flags) {
    case 0: gemm_dispatch<0>(prec, A, lda, Bt, Bhi, Blo, ldb, C, ldc, M, N, K, ep, s); break;
    case EPI_BIAS: gemm_dispatch<EPI_BIAS>(prec, A, lda, Bt, Bhi, Blo, ldb, C, ldc, M, N, K, ep, s); break;
    case EPI_V: gemm_dispatch<EPI_V>(prec, A, lda, Bt, Bhi, Blo, ldb, C, ldc, M, N, K, ep, s); break;
    case EPI_V | EPI_TANH:
      gemm_dispatch<EPI_V | EPI_TANH>(prec, A, lda, Bt, Bhi, Blo, ldb, C, ldc, M, N, K, ep, s);
      break;
    case EPI_V | EPI_TANH | EPI_PERM:
      gemm_dispatch<EPI_V | EPI_TANH | EPI_PERM>(prec, A, lda, Bt, Bhi, Blo, ldb, C, ldc, M, N, K, ep, s);
      break;
    default:
      // Remaining combinations are not used by the wavefunction; fall back to the generic instance.
      gemm_dispatch<EPI_BIAS | EPI_V | EPI_TANH | EPI_PERM>(prec, A, lda, Bt, Bhi, Blo, ldb, C, ldc, M, N, K, ep, s);
      break;
  }
}

}  // namespace dwmi

namespace dwmi {

// ------------------------------------------------------------------------------------------------
// 1-el embedding GEMM of the Laplacian pass with the tanh forward-Laplacian rule fused in the epilogue.
// Rows are (walker, electron, channel) with Cp = pad16(3 n + 2) rows per electron; a workgroup owns E
// whole electrons (RB = E * Cp / 16 row blocks of 16) and 128 output columns (4 waves x 32 columns),
// so every column of every electron has all of its channels in the tile:
//   z_ch = (X W)_ch + V[walker, ch]                     (V: spin-mean part + bias, per walker & channel)
//   y_0 = tanh z_0,  y_k = t' z_k (k = 1..3n),  y_lap = t' z_lap + t'' sum_k z_k^2,  padded channels 0.
// MFMA: v_mfma_f32_16x16x4_f32 (FP32) or v_mfma_f32_16x16x32_bf16 (BF16 forms, A split hi/lo in LDS).
// ------------------------------------------------------------------------------------------------
template <int RB, int PREC, int BN, bool VAL = false>
__global__ __launch_bounds__(256) void k_gemm_lap(const float* __restrict__ X, int ldx,
                                                  const float* __restrict__ Bt, const uint16_t* __restrict__ Bhi,
                                                  const uint16_t* __restrict__ Blo, int K,
                                                  const float* __restrict__ V, int ldv, float* __restrict__ out,
                                                  int ldo, int permute, int E, Shape sh, int tilesN, int vsplit = 1,
                                                  long vstride = 0) {
  static_assert(BN == 128 || (BN == 256 && PREC == 0), "256-column tiles exist for the FP32 form only");
  static_assert(!VAL || PREC == 0, "the value-pass variant is FP32");
  constexpr int ROWS = RB * 16;
  constexpr int WN = BN / 4;    // columns per wave
  constexpr int NQ = WN / 16;   // 16-column MFMA blocks per wave
  constexpr int BK = (PREC == 0) ? 16 : 32;  // FP32: 16 measured faster than 32 (2.05 vs 2.10 ms at N2)
  constexpr int LDK = BK + 8;  // bf16 row pitch (elements)
  // epilogue window: all 128 columns at once when the tile fits in 60 KB of LDS, else two passes of 64
  constexpr int NC = ((size_t)RB * 16 * (128 + 4) * 4 <= 60 * 1024) ? 128 : 64;
  constexpr int EP = NC + 4;   // epilogue pitch (floats)
  constexpr size_t LOAD_BYTES = (PREC == 0)
                                    ? (size_t)BK * (ROWS + 4) * 4 + (size_t)BK * (BN + 4) * 4
                                    : (size_t)(ROWS + BN) * LDK * 2 * (PREC == 1 ? 2 : 1);
  constexpr size_t EPI_BYTES = (size_t)ROWS * EP * 4;
  constexpr size_t SMEM = LOAD_BYTES > EPI_BYTES ? LOAD_BYTES : EPI_BYTES;
  __shared__ __align__(16) unsigned char smem[SMEM];

  const int Cp = sh.Cp, C = sh.C, n = sh.n;
  const long n_el_total = (long)sh.B * n;
  // XCD-aware order: the column tiles of a row tile are consecutive tiles of one XCD's range, so the row
  // tile's second read of X is served by that XCD's L2 (the grid is a multiple of 8)
  const long tiles_total = ((long)sh.B * sh.n + E - 1) / E * tilesN;
  const int tile = tile_index(blockIdx.x, (int)tiles_total);
  if (tile >= tiles_total) return;  // uniform over the workgroup, before any barrier
  const int tile_m = tile / tilesN, tile_n = tile % tilesN;
  const long g0 = (long)tile_m * E;          // first electron of the tile
  const long row0 = g0 * Cp;
  const long rows_total = n_el_total * Cp;
  const int col0 = tile_n * BN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;

  v4f acc[RB][NQ];
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < NQ; ++b) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};

  if constexpr (PREC == 0) {
    float(*As)[ROWS + 4] = (float(*)[ROWS + 4])smem;
    float(*Bs)[BN + 4] = (float(*)[BN + 4])(smem + (size_t)BK * (ROWS + 4) * 4);
    constexpr int KV = BK / 4;         // float4 per row of a K-slice
    constexpr int AV = ROWS * KV;      // float4 per A tile
    constexpr int AIT = (AV + 255) / 256;
    constexpr int BIT = BN * KV / 256;
    float4 ra[AIT], rb[BIT];
    auto gload = [&](int k0) {
#pragma unroll
      for (int it = 0; it < AIT; ++it) {
        const int idx = tid + 256 * it;
        if (idx < AV) {
          const int m = idx / KV, kq = (idx % KV) * 4;
          long gr = row0 + m;
          if (gr > rows_total - 1) gr = rows_total - 1;
          ra[it] = *(const float4*)(X + gr * ldx + k0 + kq);
        }
      }
#pragma unroll
      for (int it = 0; it < BIT; ++it) {
        const int idx = tid + 256 * it, m = idx / KV, kq = (idx % KV) * 4;
        rb[it] = *(const float4*)(Bt + (long)(col0 + m) * K + k0 + kq);
      }
    };
    auto sstore = [&]() {
#pragma unroll
      for (int it = 0; it < AIT; ++it) {
        const int idx = tid + 256 * it;
        if (idx < AV) {
          const int m = idx / KV, kq = (idx % KV) * 4;
          As[kq + 0][m] = ra[it].x;
          As[kq + 1][m] = ra[it].y;
          As[kq + 2][m] = ra[it].z;
          As[kq + 3][m] = ra[it].w;
        }
      }
#pragma unroll
      for (int it = 0; it < BIT; ++it) {
        const int idx = tid + 256 * it, m = idx / KV, kq = (idx % KV) * 4;
        Bs[kq + 0][m] = rb[it].x;
        Bs[kq + 1][m] = rb[it].y;
        Bs[kq + 2][m] = rb[it].z;
        Bs[kq + 3][m] = rb[it].w;
      }
    };
    gload(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
      __syncthreads();
      sstore();
      __syncthreads();
      if (k0 + BK < K) gload(k0 + BK);
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        // MFMA step kk / 4 of lane group g = lane / 16 takes k = kk / 4 + 4 g (A and B alike): the four rows
        // one read touches lie 4 k apart, so with the pitch of 4 (mod 64) floats they start 16 banks apart
        const int kr = (kk >> 2) + 4 * (lane >> 4);
        float bq[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) bq[q] = Bs[kr][wv * WN + q * 16 + (lane & 15)];
#pragma unroll
        for (int rb2 = 0; rb2 < RB; ++rb2) {
          const float a = As[kr][rb2 * 16 + (lane & 15)];
#pragma unroll
          for (int q = 0; q < NQ; ++q)
            acc[rb2][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bq[q], acc[rb2][q], 0, 0, 0);
        }
      }
    }
  } else {
    constexpr bool X3 = (PREC == 1);
    uint16_t(*Ah)[LDK] = (uint16_t(*)[LDK])smem;
    uint16_t(*Bh)[LDK] = (uint16_t(*)[LDK])(smem + (size_t)ROWS * LDK * 2);
    uint16_t(*Al)[LDK] = (uint16_t(*)[LDK])(smem + (size_t)(ROWS + BN) * LDK * 2);
    uint16_t(*Bl)[LDK] = (uint16_t(*)[LDK])(smem + (size_t)(2 * ROWS + BN) * LDK * 2);
    constexpr int AV = ROWS * BK / 4;
    constexpr int AIT = (AV + 255) / 256;
    float4 ra[AIT];
    uint2 rbh[4], rbl[4];
    auto gload = [&](int k0) {
#pragma unroll
      for (int it = 0; it < AIT; ++it) {
        const int idx = tid + 256 * it;
        if (idx < AV) {
          const int m = idx >> 3, kq = (idx & 7) * 4;
          long gr = row0 + m;
          if (gr > rows_total - 1) gr = rows_total - 1;
          ra[it] = *(const float4*)(X + gr * ldx + k0 + kq);
        }
      }
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + 256 * it, m = idx >> 3, kq = (idx & 7) * 4;
        rbh[it] = *(const uint2*)(Bhi + (long)(col0 + m) * K + k0 + kq);
        if (X3) rbl[it] = *(const uint2*)(Blo + (long)(col0 + m) * K + k0 + kq);
      }
    };
    auto sstore = [&]() {
#pragma unroll
      for (int it = 0; it < AIT; ++it) {
        const int idx = tid + 256 * it;
        if (idx < AV) {
          const int m = idx >> 3, kq = (idx & 7) * 4;
          const float f[4] = {ra[it].x, ra[it].y, ra[it].z, ra[it].w};
          unsigned short h[4], l[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            h[q] = f2bf_bits(f[q]);
            l[q] = f2bf_bits(f[q] - bf_bits2f(h[q]));
          }
          uint2 hv, lv;
          hv.x = h[0] | ((unsigned)h[1] << 16);
          hv.y = h[2] | ((unsigned)h[3] << 16);
          lv.x = l[0] | ((unsigned)l[1] << 16);
          lv.y = l[2] | ((unsigned)l[3] << 16);
          *(uint2*)&Ah[m][kq] = hv;
          if (X3) *(uint2*)&Al[m][kq] = lv;
        }
      }
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int idx = tid + 256 * it, m = idx >> 3, kq = (idx & 7) * 4;
        *(uint2*)&Bh[m][kq] = rbh[it];
        if (X3) *(uint2*)&Bl[m][kq] = rbl[it];
      }
    };
    gload(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
      __syncthreads();
      sstore();
      __syncthreads();
      if (k0 + BK < K) gload(k0 + BK);
      const int kc = 8 * (lane >> 4);
      v8s bh[2], bl[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        bh[q] = *(const v8s*)&Bh[wv * 32 + q * 16 + (lane & 15)][kc];
        if (X3) bl[q] = *(const v8s*)&Bl[wv * 32 + q * 16 + (lane & 15)][kc];
      }
#pragma unroll
      for (int rb2 = 0; rb2 < RB; ++rb2) {
        const v8s ah = *(const v8s*)&Ah[rb2 * 16 + (lane & 15)][kc];
        v8s al;
        if (X3) al = *(const v8s*)&Al[rb2 * 16 + (lane & 15)][kc];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if (X3) {
            acc[rb2][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16((v8bf)al, (v8bf)bh[q], acc[rb2][q], 0, 0, 0);
            acc[rb2][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16((v8bf)ah, (v8bf)bl[q], acc[rb2][q], 0, 0, 0);
          }
          acc[rb2][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16((v8bf)ah, (v8bf)bh[q], acc[rb2][q], 0, 0, 0);
        }
      }
    }
  }
  if constexpr (VAL) {
    // value pass (one row per electron): y = tanh(acc + V[walker]) straight from the accumulators, V the
    // sum of its K-split parts (bias included), rows optionally to spin-major order
    const int dn = n - sh.u;
#pragma unroll
    for (int rb2 = 0; rb2 < RB; ++rb2)
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const long r = row0 + rb2 * 16 + 4 * (lane >> 4) + j;
          if (r < rows_total) {
            const int c = col0 + wv * WN + q * 16 + (lane & 15);
            const long b = div_idx(r, n);
            const int i = (int)(r - b * n);
            const float* Vp = V + b * (long)ldv + c;
            float vs = Vp[0];
            for (int t = 1; t < vsplit; ++t) vs += Vp[t * vstride];
            const float y = tanh_fast(acc[rb2][q][j] + vs);
            long orow = r;
            if (permute) orow = (i < sh.u) ? b * sh.u + i : (long)sh.B * sh.u + b * dn + (i - sh.u);
            out[orow * ldo + c] = y;
          }
        }
    return;
  }
  // ---- epilogue: tile to LDS, then (electron, column) pairs with lanes along columns, so every store of
  // a wave is one contiguous 256-byte row segment.  When a pass has fewer than 256 pairs, G threads share
  // a pair and split its tangent channels into contiguous chunks; their partial sum_k z_k^2 meet in LDS
  // in a fixed order (deterministic).  V and z of 8 channels are loaded before any of them is used.
  __shared__ float red[256];
  float(*Es)[EP] = (float(*)[EP])smem;
  const int dn = n - sh.u;
  const int P = E * NC;
  const int G = P >= 256 ? 1 : 256 / P;
  const int nt = C - 2;
  const int chunk = (nt + G - 1) / G;
#pragma unroll
  for (int pass = 0; pass < BN / NC; ++pass) {
    __syncthreads();
    const int wc0 = wv * WN - pass * NC;  // this wave's first column inside the pass window
    if (wc0 >= 0 && wc0 < NC) {
#pragma unroll
      for (int rb2 = 0; rb2 < RB; ++rb2)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
          for (int j = 0; j < 4; ++j) Es[rb2 * 16 + 4 * (lane >> 4) + j][wc0 + q * 16 + (lane & 15)] = acc[rb2][q][j];
    }
    __syncthreads();
    for (int t0 = 0; t0 < P * G; t0 += 256) {
      const int t = t0 + tid;
      const int grp = t / P, p = t - (t / P) * P;
      const int e = p / NC, c = p - (p / NC) * NC;
      const long g = g0 + e;
      const bool valid = (t < P * G) && (g < n_el_total);
      const int gc = col0 + pass * NC + c;
      long obase = 0;
      const float* Vb = V;
      const float* Ez = &Es[0][0];
      float d1 = 0.f, d2 = 0.f, S = 0.f;
      if (valid) {
        const long b = g / n;
        const int i = (int)(g - b * n);
        Vb = V + (b * Cp) * (long)ldv + gc;
        if (permute)
          obase = (i < sh.u) ? (b * sh.u + i) * Cp : (long)sh.B * sh.u * Cp + (b * dn + (i - sh.u)) * Cp;
        else
          obase = g * Cp;
        Ez = &Es[e * Cp][c];
        const float th = tanh_fast(Ez[0] + Vb[0]);
        d1 = 1.f - th * th;
        d2 = -2.f * th * d1;
        if (grp == 0) out[obase * ldo + gc] = th;
        const int c_lo = 1 + grp * chunk;
        const int c_hi = min(C - 1, c_lo + chunk);
        int ch = c_lo;
        for (; ch + 8 <= c_hi; ch += 8) {
          float z[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) z[u] = Ez[(ch + u) * EP] + Vb[(long)(ch + u) * ldv];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            S += z[u] * z[u];
            out[(obase + ch + u) * ldo + gc] = d1 * z[u];
          }
        }
        for (; ch < c_hi; ++ch) {
          const float zc = Ez[ch * EP] + Vb[(long)ch * ldv];
          S += zc * zc;
          out[(obase + ch) * ldo + gc] = d1 * zc;
        }
      }
      if (G > 1) {
        red[tid] = S;
        __syncthreads();
        S = 0.f;
        for (int q = 0; q < G; ++q) S += red[q * P + p];
      }
      if (valid && grp == 0) {
        const float zl = Ez[(C - 1) * EP] + Vb[(long)(C - 1) * ldv];
        out[(obase + C - 1) * ldo + gc] = d1 * zl + d2 * S;
        for (int ch = C; ch < Cp; ++ch) out[(obase + ch) * ldo + gc] = 0.f;
      }
    }
  }
}

template <int PREC>
static void gemm_lap_rb(int RB, const float* X, int ldx, const float* Bt, const uint16_t* Bhi, const uint16_t* Blo,
                        int K, const float* V, int ldv, float* out, int ldo, int permute, int E, const Shape& sh,
                        int cols, hipStream_t s) {
  const long tilesM = ((long)sh.B * sh.n + E - 1) / E;
  if constexpr (PREC == 0) {
    // 256-column tiles (every output column of a row tile in one workgroup, so its A tile is read once)
    if (cols == 256 && RB <= 6 && sh.W1 % 256 == 0) {
      const int tilesN = sh.W1 / 256;
      const dim3 grid((unsigned)((tilesM * tilesN + 7) / 8 * 8));
#define DW_LAP_CASE256(R)                                                                                            \
  case R:                                                                                                            \
    hipLaunchKernelGGL((k_gemm_lap<R, 0, 256>), grid, dim3(256), 0, s, X, ldx, Bt, Bhi, Blo, K, V, ldv, out, ldo, \
                       permute, E, sh, tilesN);                                                                      \
    return;
      switch (RB) {
        DW_LAP_CASE256(1)
        DW_LAP_CASE256(2)
        DW_LAP_CASE256(3)
        DW_LAP_CASE256(4)
        DW_LAP_CASE256(5)
        DW_LAP_CASE256(6)
        default: break;
      }
#undef DW_LAP_CASE256
    }
  }
  const int tilesN = sh.W1 / 128;
  const dim3 grid((unsigned)((tilesM * tilesN + 7) / 8 * 8));
#define DW_LAP_CASE(R)                                                                                                  \
  case R:                                                                                                               \
    hipLaunchKernelGGL((k_gemm_lap<R, PREC, 128>), grid, dim3(256), 0, s, X, ldx, Bt, Bhi, Blo, K, V, ldv, out, ldo, \
                       permute, E, sh, tilesN);                                                                         \
    break;
  switch (RB) {
    DW_LAP_CASE(1)
    DW_LAP_CASE(2)
    DW_LAP_CASE(3)
    DW_LAP_CASE(4)
    DW_LAP_CASE(5)
    DW_LAP_CASE(6)
    DW_LAP_CASE(7)
    DW_LAP_CASE(8)
    DW_LAP_CASE(9)
    DW_LAP_CASE(10)
    DW_LAP_CASE(11)
    DW_LAP_CASE(12)
    default: break;
  }
#undef DW_LAP_CASE
}

void gemm_value16(const float* X, int ldx, const float* Bt, int K, const float* V, int ldv, int vsplit, long vstride,
                  float* out, int ldo, int permute, const Shape& sh, hipStream_t s) {
  // 96-electron x 128-column tiles on 16x16x4 MFMA (the Laplacian GEMM's main loop), one row per electron
  Shape v = sh;
  v.C = 1;
  v.Cp = 1;
  const int E = 96;
  const int tilesN = sh.W1 / 128;
  const long tilesM = ((long)sh.B * sh.n + E - 1) / E;
  const dim3 grid((unsigned)((tilesM * tilesN + 7) / 8 * 8));
  hipLaunchKernelGGL((k_gemm_lap<6, 0, 128, true>), grid, dim3(256), 0, s, X, ldx, Bt,
                     nullptr, nullptr, K, V, ldv, out, ldo, permute, E, v, tilesN, vsplit, vstride);
}

// ------------------------------------------------------------------------------------------------
// Per-walker term of a 1-el layer in the Laplacian pass (FP32), spin means fused in:
//   V[r, o] = sum_k MU[r, k] Wmu[o, k] (+ bias[o] on each walker's value channel),
//   MU[r, k] = mean over the electrons of spin s(k) of X[(b n + i) Cp + ch, k - s Kh],  r = b Cp + ch,
// with MU formed in LDS for 48 rows and 128 of its columns at a time (never stored; 25 KB of LDS and 124
// registers, so four workgroups per CU overlap one another's load and MFMA phases: 522 -> 478 us per N2 layer).  The means add the
// electrons in order and scale by 1 / n_s like k_spin_mean, and the MFMA sequence per output element
// (16-wide K slices in order, step s of lane group g taking k = 4 g + s) is k_gemm_f32's, so V equals
// what those two kernels produce.  4 waves x 64 output columns (N = 256); each wave's weights come
// straight from global memory (L2).
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mean_gemm(const float* __restrict__ X, int Kx, int Kh,
                                                   const float* __restrict__ Bt, int Km, const float* __restrict__ bias,
                                                   float* __restrict__ V, int ldv, Shape sh) {
  constexpr int RT = 48, RB = RT / 16, CW = 128, PITCH = CW + 4;
  __shared__ __align__(16) float Ms[RT][PITCH];
  const long rows = (long)sh.B * sh.Cp;
  const long r0 = (long)blockIdx.x * RT;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, cl = lane & 15;
  const int cw = wv * 64;  // this wave's first output column
  const int n = sh.n, u = sh.u, Cp = sh.Cp;
  const long xstride = (long)Cp * Kx;  // one electron further

  v4f acc[RB][4];
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[a][q] = v4f{0.f, 0.f, 0.f, 0.f};

  for (int kc = 0; kc < Km; kc += CW) {
    const int cwid = Km - kc < CW ? Km - kc : CW;  // a multiple of 16
    const int q4 = cwid / 4;
    __syncthreads();  // the previous chunk's reads of Ms are done
    for (int e = tid; e < RT * q4; e += 256) {
      const int rr = e / q4, f = (e - rr * q4) * 4;
      const long r = r0 + rr;
      const int k = kc + f;
      float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f;
      if (r < rows && k < 2 * Kh) {
        const long b = div_idx(r, Cp);
        const int ch = (int)(r - b * Cp);
        const int s = k >= Kh ? 1 : 0;
        const int i0 = s ? u : 0, i1 = s ? n : u;
        const float* xp = X + ((b * n + i0) * Cp + ch) * (long)Kx + (k - s * Kh);
#pragma unroll 4
        for (int i = i0; i < i1; ++i, xp += xstride) {
          const v4f x = *(const v4f*)xp;
          v0 += x[0];
          v1 += x[1];
          v2 += x[2];
          v3 += x[3];
        }
        const float inv = 1.f / (float)(i1 - i0);
        v0 *= inv;
        v1 *= inv;
        v2 *= inv;
        v3 *= inv;
      }
      *(v4f*)&Ms[rr][f] = v4f{v0, v1, v2, v3};
    }
    __syncthreads();
    v4f bq[4];
    auto bload = [&](int k0, v4f* b) {
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = *(const v4f*)(Bt + (long)(cw + 16 * q + cl) * Km + kc + k0 + 4 * g);
    };
    for (int k0 = 0; k0 < cwid; k0 += 16) {
      bload(k0, bq);
      v4f a4[RB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) a4[rb] = *(const v4f*)&Ms[rb * 16 + cl][k0 + 4 * g];
#pragma unroll
      for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            acc[rb][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[rb][st], bq[q][st], acc[rb][q], 0, 0, 0);
    }
  }
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const long r = r0 + rb * 16 + 4 * g + e;
        if (r < rows) {
          const int c = cw + 16 * q + cl;
          float v = acc[rb][q][e];
          if (bias && r - div_idx(r, Cp) * Cp == 0) v += bias[c];
          V[r * ldv + c] = v;
        }
      }
}

bool gemm_mean_lap(const float* X, int Kx, int Kh, const float* Wmu, int Km, const float* bias, float* V, int ldv,
                   const Shape& sh, hipStream_t s) {
  // (the layer-0 GEMM, K of a few features, runs faster as the two separate kernels)
  if (sh.W1 != 256 || Km < 256 || Km % 16 || Kh % 4 || Kx % 4 || sh.C < 2) return false;
  const long rows = (long)sh.B * sh.Cp;
  hipLaunchKernelGGL(k_mean_gemm, dim3((unsigned)((rows + 47) / 48)), dim3(256), 0, s, X, Kx, Kh, Wmu, Km, bias, V,
                     ldv, sh);
  return true;
}

int lap_electrons_per_tile(int Cp, GemmPrec prec) {
  // FP32: the most electrons (at most 4) whose rows fit 192; measured 1.7-2.4 % faster than 96 rows at
  // 4, 14, 17 and 26 electrons.  BF16 forms keep 96 rows.
  int E = prec == GEMM_F32 ? 192 / Cp : 96 / Cp;
  if (prec == GEMM_F32 && E > 4) E = 4;
  return E < 1 ? 1 : E;
}

void gemm_lap(GemmPrec prec, const float* X, int ldx, const float* Bt, const uint16_t* Bhi, const uint16_t* Blo, int K,
              const float* V, int ldv, float* out, int ldo, int permute, const Shape& sh, int E_override, int cols,
              hipStream_t s) {
  const int E = E_override > 0 ? E_override : lap_electrons_per_tile(sh.Cp, prec);
  const int RB = E * sh.Cp / 16;
  if (prec == GEMM_F32)
    gemm_lap_rb<0>(RB, X, ldx, Bt, Bhi, Blo, K, V, ldv, out, ldo, permute, E, sh, cols, s);
  else if (prec == GEMM_BF16X3)
    gemm_lap_rb<1>(RB, X, ldx, Bt, Bhi, Blo, K, V, ldv, out, ldo, permute, E, sh, cols, s);
  else
    gemm_lap_rb<2>(RB, X, ldx, Bt, Bhi, Blo, K, V, ldv, out, ldo, permute, E, sh, cols, s);
}

}  // namespace dwmi
