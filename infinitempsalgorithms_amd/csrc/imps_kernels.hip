// gfx950 kernels of libimps: MFMA GEMM for the bulk contractions, the MPO mixing
// kernels, FP64 reductions and the small bandwidth-bound helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>

#include "imps_internal.h"

namespace imps {

typedef double v4d __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------
// GEMM on MFMA, FP64 data in HBM.
//
// TC = double: v_mfma_f64_16x16x4_f64 (FP64 products and accumulation).
// TC = float : operands rounded to FP32 when staged into LDS, v_mfma_f32_16x16x4_f32
//              (FP32 products and accumulation), result widened to FP64 in the epilogue.
//
// Block = 4 wavefronts (2 x 2), macro tile BM x BN, K step BK.  Operands are staged
// global -> registers -> LDS (k-major rows, one padding element per row); the next
// K tile is fetched into registers while the MFMAs consume the current one.
// The MFMA is issued with the B fragment as its row operand so that the 16 lanes of
// an accumulator row hold 16 consecutive m of the column-major C (coalesced stores).
// Accumulator layouts (measured): f64 16x16x4 register r of lane l holds row l/16 + 4r;
// f32 16x16x4 holds row 4(l/16) + r; both hold column l % 16.
// ---------------------------------------------------------------------------------

template <typename TC>
struct MfmaTraits;
template <>
struct MfmaTraits<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct MfmaTraits<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lk, int r) { return 4 * lk + r; }
};

// Output tile of workgroup (bx, by) in an MT x NT grid.  GM = 0: tile (bx, by).  GM > 0: the
// hardware hands consecutive workgroup ids to the 8 XCDs in turn, so id i runs on XCD i % 8; the
// tiles are renumbered so that each XCD owns a contiguous range of tile numbers (when 8 divides
// MT NT), and numbers run over groups of GM tile rows, m fastest inside a group, so that the
// workgroups resident on one XCD share A and B panels in its L2.  A bijection of the tiles, so
// every output is computed exactly as in the identity order.
template <int GM>
__device__ inline void gemm_tile_of(int bx, int by, int MT, int NT, int& mt, int& nt) {
  if (GM == 0) {
    mt = bx;
    nt = by;
    return;
  }
  const int T = MT * NT;
  const int lin = bx + MT * by;
  const int L = (T % 8 == 0) ? (lin % 8) * (T / 8) + lin / 8 : lin;
  const int per = GM * NT;
  const int grp = L / per;
  const int first = grp * GM;
  const int gsz = (MT - first < GM) ? MT - first : GM;
  const int r = L - grp * per;
  mt = first + r % gsz;
  nt = r / gsz;
}

// Epilogue of every tile variant: C = fma(beta, C_old, alpha * acc), one rounding order for all of
// them, so that no variant depends on how the compiler contracts alpha acc + beta C_old.
// VEC: every tile is full (M % BM, N % BN, K % BK, Kc % BK all zero) and each thread moves
// 16-byte pairs along the contiguous direction of A and B (even strides, 16-byte aligned bases),
// so the tile loads need no bounds checks.  GM: tile order (gemm_tile_of).
template <typename TC, int BM, int BN, int BK, bool AMC, bool BKC, int OCC, bool VEC, int GM = 0, bool PRIO = false>
__global__ void __launch_bounds__(256, OCC) gemm_kernel(GemmDesc g) {
  typedef MfmaTraits<TC> MT;
  typedef typename MT::v4 v4;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int AE = BM * BK / 256, BE = BK * BN / 256;
  static_assert(AE * 256 == BM * BK && BE * 256 == BK * BN, "tile/thread mismatch");

  __shared__ TC As[BK][BM + 1];
  __shared__ TC Bs[BK][BN + 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  int mt, nt;
  gemm_tile_of<GM>(blockIdx.x, blockIdx.y, gridDim.x, gridDim.y, mt, nt);
  const int m0 = mt * BM;
  const int n0 = nt * BN;
  const int wm = (wave & 1) * WM;
  const int wn = (wave >> 1) * WN;
  const int M = g.M, N = g.N, K = g.K, Kc = g.Kc;
  // split-K: blockIdx.z owns K tiles [kt0, kt1)
  const int ntiles_all = (K + BK - 1) / BK;
  const int per = (ntiles_all + gridDim.z - 1) / gridDim.z;
  const int kt0 = blockIdx.z * per;
  const int kt1 = min(ntiles_all, kt0 + per);

  v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4{0, 0, 0, 0};

  double ra[AE], rb[BE];
  typedef double v2d __attribute__((ext_vector_type(2)));
  static_assert(!VEC || (AE % 2 == 0 && BE % 2 == 0), "paired loads need an even count");

  // tile element e of this thread: (m, k) of A and (k, n) of B
  auto a_index = [&](int e, int& m, int& k) {
    const int idx = VEC ? tid + 256 * (e / 2) : tid + 256 * e;
    if (AMC) {
      m = VEC ? (idx % (BM / 2)) * 2 + (e & 1) : idx % BM;
      k = VEC ? idx / (BM / 2) : idx / BM;
    } else {
      k = VEC ? (idx % (BK / 2)) * 2 + (e & 1) : idx % BK;
      m = VEC ? idx / (BK / 2) : idx / BK;
    }
  };
  auto b_index = [&](int e, int& k, int& n) {
    const int idx = VEC ? tid + 256 * (e / 2) : tid + 256 * e;
    if (BKC) {
      k = VEC ? (idx % (BK / 2)) * 2 + (e & 1) : idx % BK;
      n = VEC ? idx / (BK / 2) : idx / BK;
    } else {
      n = VEC ? (idx % (BN / 2)) * 2 + (e & 1) : idx % BN;
      k = VEC ? idx / (BN / 2) : idx / BN;
    }
  };

  // Stage bases (wave-uniform): K chunk kc and offset kk inside it, advanced by BK per stage
  // (no division in the loop); VEC loads add per-thread byte offsets computed once (they fit
  // in 31 bits: the host checks the extent of the operands).
  int kc = (kt0 * BK) / Kc;
  int kk = kt0 * BK - kc * Kc;
  const double* Ab = g.A + (long)kc * g.sAc + (long)kk * g.sAk;
  const double* Bb = g.B + (long)kc * g.sBc + (long)kk * g.sBk;
  auto advance = [&]() {
    kk += BK;
    if (kk >= Kc) {
      kk -= Kc;
      ++kc;
      Ab = g.A + (long)kc * g.sAc + (long)kk * g.sAk;
      Bb = g.B + (long)kc * g.sBc + (long)kk * g.sBk;
    } else {
      Ab += (long)BK * g.sAk;
      Bb += (long)BK * g.sBk;
    }
  };
  // byte offsets (unsigned 32-bit): a wave-uniform base plus one VGPR per load
  unsigned offA[VEC ? AE / 2 : 1], offB[VEC ? BE / 2 : 1];
  if (VEC) {
#pragma unroll
    for (int e = 0; e < AE; e += 2) {
      int m, k;
      a_index(e, m, k);
      offA[e / 2] = (unsigned)(8 * ((long)(m0 + m) * g.sAm + (long)k * g.sAk));
    }
#pragma unroll
    for (int e = 0; e < BE; e += 2) {
      int k, n;
      b_index(e, k, n);
      offB[e / 2] = (unsigned)(8 * ((long)k * g.sBk + (long)(n0 + n) * g.sBn));
    }
  }

  auto load_tiles = [&](int k0) {
    if (VEC) {
#pragma unroll
      for (int e = 0; e < AE; e += 2) {
        const v2d v = *(const v2d*)((const char*)Ab + offA[e / 2]);
        ra[e] = v[0];
        ra[e + 1] = v[1];
      }
#pragma unroll
      for (int e = 0; e < BE; e += 2) {
        const v2d v = *(const v2d*)((const char*)Bb + offB[e / 2]);
        rb[e] = v[0];
        rb[e + 1] = v[1];
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      int m, k;
      a_index(e, m, k);
      const int gm = m0 + m;
      ra[e] = (gm < M && k0 + k < K) ? Ab[(long)gm * g.sAm + (long)k * g.sAk] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      int k, n;
      b_index(e, k, n);
      const int gn = n0 + n;
      rb[e] = (gn < N && k0 + k < K) ? Bb[(long)k * g.sBk + (long)gn * g.sBn] : 0.0;
    }
  };

  auto store_tiles = [&]() {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      int m, k;
      a_index(e, m, k);
      As[k][m] = (TC)ra[e];
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      int k, n;
      b_index(e, k, n);
      Bs[k][n] = (TC)rb[e];
    }
  };

  if (kt0 < kt1) load_tiles(kt0 * BK);
  const int lr = lane & 15;
  const int lk = lane >> 4;
  for (int kt = kt0; kt < kt1; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < kt1) {
      advance();
      load_tiles((kt + 1) * BK);
    }
    if (PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int k4 = 0; k4 < BK; k4 += 4) {
      TC af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[k4 + lk][wm + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[k4 + lk][wn + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = MT::mma(bf[j], af[i], acc[i][j]);
    }
    if (PRIO) __builtin_amdgcn_s_setprio(0);
  }

  // split-K partial results go to the workspace slice of this z; the reduction kernel
  // applies alpha and beta
  double* Cout = g.C;
  long ldc = g.ldc;
  double alpha = g.alpha, beta = g.beta;
  if (gridDim.z > 1) {
    Cout = g.W + (size_t)blockIdx.z * M * N;
    ldc = M;
    alpha = 1.0;
    beta = 0.0;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int m = m0 + wm + i * 16 + lr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + j * 16 + MT::row(lk, r);
        if (n < N) {
          double* p = Cout + m + (long)n * ldc;
          const double v = alpha * (double)acc[i][j][r];
          *p = (beta == 0.0) ? v : __builtin_fma(beta, *p, v);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------
// Split-BF16 three-term GEMM: every FP64 operand x is staged as x_hi = bf16(x) and
// x_lo = bf16(x - x_hi) (|x - x_hi - x_lo| <~ 2^-17 |x|); a*b is accumulated in FP32 as
// a_hi b_hi + a_hi b_lo + a_lo b_hi with v_mfma_f32_16x16x32_bf16.  LDS rows are k-contiguous
// (BK = 32 bf16 plus 8 of padding) so each lane fetches its 8 k-values with one 16-byte read.
// ---------------------------------------------------------------------------------

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline void split_bf16(double x, __bf16& hi, __bf16& lo) {
  const float xf = (float)x;
  hi = (__bf16)xf;
  lo = (__bf16)(xf - (float)hi);
}

template <int BM, int BN, bool AMC, bool BKC>
__global__ void __launch_bounds__(256) gemm_bf16x3_kernel(GemmDesc g) {
  constexpr int BK = 32, LDK = BK + 8;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int AE = BM * BK / 256, BE = BK * BN / 256;
  __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Al[BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[BN][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[BN][LDK];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int m0 = blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  const int wm = (wave & 1) * WM;
  const int wn = (wave >> 1) * WN;
  const int M = g.M, N = g.N, K = g.K, Kc = g.Kc;
  const int ntiles_all = (K + BK - 1) / BK;
  const int per = (ntiles_all + gridDim.z - 1) / gridDim.z;
  const int kt0 = blockIdx.z * per;
  const int kt1 = min(ntiles_all, kt0 + per);
  v4f acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  double ra[AE], rb[BE];
  auto load_tiles = [&](int k0) {
    const int kc = k0 / Kc;
    const int kk0 = k0 - kc * Kc;
    const double* Ab = g.A + (long)kc * g.sAc + (long)kk0 * g.sAk;
    const double* Bb = g.B + (long)kc * g.sBc + (long)kk0 * g.sBk;
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int idx = tid + 256 * e;
      int m, k;
      if (AMC) {
        m = idx % BM;
        k = idx / BM;
      } else {
        k = idx % BK;
        m = idx / BK;
      }
      const int gm = m0 + m;
      ra[e] = (gm < M && k0 + k < K) ? Ab[(long)gm * g.sAm + (long)k * g.sAk] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int idx = tid + 256 * e;
      int n, k;
      if (BKC) {
        k = idx % BK;
        n = idx / BK;
      } else {
        n = idx % BN;
        k = idx / BN;
      }
      const int gn = n0 + n;
      rb[e] = (gn < N && k0 + k < K) ? Bb[(long)k * g.sBk + (long)gn * g.sBn] : 0.0;
    }
  };
  auto store_tiles = [&]() {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int idx = tid + 256 * e;
      int m, k;
      if (AMC) {
        m = idx % BM;
        k = idx / BM;
      } else {
        k = idx % BK;
        m = idx / BK;
      }
      split_bf16(ra[e], Ah[m][k], Al[m][k]);
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int idx = tid + 256 * e;
      int n, k;
      if (BKC) {
        k = idx % BK;
        n = idx / BK;
      } else {
        n = idx % BN;
        k = idx / BN;
      }
      split_bf16(rb[e], Bh[n][k], Bl[n][k]);
    }
  };
  if (kt0 < kt1) load_tiles(kt0 * BK);
  const int lr = lane & 15;
  const int lk = lane >> 4;
  for (int kt = kt0; kt < kt1; ++kt) {
    __syncthreads();
    store_tiles();
    __syncthreads();
    if (kt + 1 < kt1) load_tiles((kt + 1) * BK);
    v8bf ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      ah[i] = *(const v8bf*)&Ah[wm + i * 16 + lr][8 * lk];
      al[i] = *(const v8bf*)&Al[wm + i * 16 + lr][8 * lk];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      bh[j] = *(const v8bf*)&Bh[wn + j * 16 + lr][8 * lk];
      bl[j] = *(const v8bf*)&Bl[wn + j * 16 + lr][8 * lk];
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
      }
  }
  double* Cout = g.C;
  long ldc = g.ldc;
  double alpha = g.alpha, beta = g.beta;
  if (gridDim.z > 1) {
    Cout = g.W + (size_t)blockIdx.z * M * N;
    ldc = M;
    alpha = 1.0;
    beta = 0.0;
  }
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int m = m0 + wm + i * 16 + lr;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + j * 16 + 4 * lk + r;
        if (n < N) {
          double* p = Cout + m + (long)n * ldc;
          const double v = alpha * (double)acc[i][j][r];
          *p = (beta == 0.0) ? v : __builtin_fma(beta, *p, v);
        }
      }
    }
  }
}

// C = alpha * sum_z W[z] + beta * C
__global__ void splitk_reduce_kernel(int M, int N, int nz, const double* __restrict__ W, double alpha,
                                     double beta, double* C, long ldc) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (m >= M) return;
  double s = 0.0;
  for (int z = 0; z < nz; ++z) s += W[(size_t)z * M * N + m + (size_t)M * n];
  double* p = C + m + ldc * n;
  *p = (beta == 0.0) ? alpha * s : __builtin_fma(beta, *p, alpha * s);
}

template <typename TC, int BM, int BN, int BK, int OCC, bool VEC, int GM = 0, bool PRIO = false>
static hipError_t gemm_dispatch_tile(const GemmDesc& g, bool amc, bool bkc, int splitk, hipStream_t s) {
  dim3 grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, splitk);
  dim3 block(256);
  if (amc && bkc)
    hipLaunchKernelGGL((gemm_kernel<TC, BM, BN, BK, true, true, OCC, VEC, GM, PRIO>), grid, block, 0, s, g);
  else if (amc && !bkc)
    hipLaunchKernelGGL((gemm_kernel<TC, BM, BN, BK, true, false, OCC, VEC, GM, PRIO>), grid, block, 0, s, g);
  else if (!amc && bkc)
    hipLaunchKernelGGL((gemm_kernel<TC, BM, BN, BK, false, true, OCC, VEC, GM, PRIO>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<TC, BM, BN, BK, false, false, OCC, VEC, GM, PRIO>), grid, block, 0, s, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || splitk == 1) return e;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((g.M + 255) / 256, g.N), dim3(256), 0, s, g.M, g.N, splitk,
                     g.W, g.alpha, g.beta, g.C, g.ldc);
  return hipGetLastError();
}

template <int BM, int BN>
static hipError_t gemm_bf16x3_tile(const GemmDesc& g, bool amc, bool bkc, int splitk, hipStream_t s) {
  dim3 grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, splitk);
  dim3 block(256);
  if (amc && bkc)
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, BN, true, true>), grid, block, 0, s, g);
  else if (amc && !bkc)
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, BN, true, false>), grid, block, 0, s, g);
  else if (!amc && bkc)
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, BN, false, true>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, BN, false, false>), grid, block, 0, s, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || splitk == 1) return e;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((g.M + 255) / 256, g.N), dim3(256), 0, s, g.M, g.N, splitk,
                     g.W, g.alpha, g.beta, g.C, g.ldc);
  return hipGetLastError();
}

size_t gemm_workspace(int M, int N, int K) {
  // doubles needed by the split-K choice of gemm_select (0: no split-K for this shape)
  (void)K;
  const long t128 = (long)((M + 127) / 128) * ((N + 127) / 128);
  const long t64 = (long)((M + 63) / 64) * ((N + 63) / 64);
  if (t128 >= 192 || t64 >= 512) return 0;
  return (size_t)8 * M * N;
}

// true when every BM x BN x 16 tile is full, the paired 16-byte loads stay aligned and the
// per-thread byte offsets inside one K stage fit in 31 bits
static bool gemm_vec_ok(const GemmDesc& g, bool amc, bool bkc, int BM, int BN) {
  auto even = [](long x) { return (x & 1) == 0; };
  if (g.M % BM || g.N % BN || g.K % 16 || g.Kc % 16) return false;
  const long extA = (long)g.M * std::labs(g.sAm) + 16L * std::labs(g.sAk);
  const long extB = 16L * std::labs(g.sBk) + (long)g.N * std::labs(g.sBn);
  if (extA >= (1L << 28) || extB >= (1L << 28)) return false;
  if (((uintptr_t)g.A & 15) || ((uintptr_t)g.B & 15)) return false;
  const bool a_ok = amc ? (even(g.sAk) && even(g.sAc)) : (even(g.sAm) && even(g.sAc));
  const bool b_ok = bkc ? (even(g.sBn) && even(g.sBc)) : (even(g.sBk) && even(g.sBc));
  return a_ok && b_ok;
}

template <typename TC>
static hipError_t gemm_select(const GemmDesc& g, bool amc, bool bkc, hipStream_t s) {
  const long t64x128 = (long)((g.M + 63) / 64) * ((g.N + 127) / 128);
  // large shapes: 64 x 128 tiles at four workgroups per CU with paired loads (variant 5)
  if (t64x128 >= 384 && gemm_vec_ok(g, amc, bkc, 64, 128))
    return gemm_dispatch_tile<TC, 64, 128, 16, 4, true, 8, true>(g, amc, bkc, 1, s);
  const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
  if (t128 >= 192) return gemm_dispatch_tile<TC, 128, 128, 16, 2, false>(g, amc, bkc, 1, s);
  const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
  int splitk = 1;
  if (g.W != nullptr && t64 < 512) {
    // long-K shapes: split K so that about 1024 workgroups run; each split keeps >= 256 K
    const int ktiles = (g.K + 15) / 16;
    splitk = (int)std::min<long>(8, std::max<long>(1, 1024 / std::max<long>(t64, 1)));
    splitk = std::min(splitk, std::max(1, ktiles / 16));
  }
  if (gemm_vec_ok(g, amc, bkc, 64, 64))
    return gemm_dispatch_tile<TC, 64, 64, 16, 4, true>(g, amc, bkc, splitk, s);
  return gemm_dispatch_tile<TC, 64, 64, 16, 2, false>(g, amc, bkc, splitk, s);
}

// split-K depth gemm_select takes for this shape (1: none)
static int gemm_select_splitk(const GemmDesc& g, bool amc, bool bkc) {
  const long t64x128 = (long)((g.M + 63) / 64) * ((g.N + 127) / 128);
  if (t64x128 >= 384 && gemm_vec_ok(g, amc, bkc, 64, 128)) return 1;
  const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
  if (t128 >= 192) return 1;
  const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
  if (g.W == nullptr || t64 >= 512) return 1;
  const int ktiles = (g.K + 15) / 16;
  int splitk = (int)std::min<long>(8, std::max<long>(1, 1024 / std::max<long>(t64, 1)));
  return std::min(splitk, std::max(1, ktiles / 16));
}

// checks the descriptor and normalises Kc; false when it is invalid
static bool gemm_normalise(GemmDesc& g, bool& amc, bool& bkc) {
  if (g.Kc <= 0 || g.Kc >= g.K) g.Kc = std::max(g.K, 1);
  if (g.Kc < g.K && (g.Kc % 16) != 0) return false;
  amc = (g.sAm == 1);
  bkc = (g.sBk == 1);
  if (!amc && g.sAk != 1) return false;
  if (!bkc && g.sBn != 1) return false;
  return true;
}

// Tile variants of the FP64/FP32 GEMM without split-K.  They accumulate every output over K in
// the same order (K tiles of 16, MFMA steps of 4), so their results are bit-identical and the
// choice among them is a matter of speed only:
//   0: 64 x 128, occupancy 4, paired loads, XCD-grouped tile order (groups of 8 tile rows)
//   1: 64 x 64, occupancy 4, paired loads   4: variant 0 in the identity tile order
//   2: 128 x 128, occupancy 2                3: 64 x 64, occupancy 2
//   5: variant 0 with the wave priority raised while it issues MFMAs (s_setprio), so that a wave
//      with MFMAs ready wins the issue arbitration over one storing its next LDS stage
template <typename TC>
static hipError_t gemm_variant(const GemmDesc& g, bool amc, bool bkc, int v, hipStream_t s) {
  switch (v) {
    case 5: return gemm_dispatch_tile<TC, 64, 128, 16, 4, true, 8, true>(g, amc, bkc, 1, s);
    case 0: return gemm_dispatch_tile<TC, 64, 128, 16, 4, true, 8>(g, amc, bkc, 1, s);
    case 1: return gemm_dispatch_tile<TC, 64, 64, 16, 4, true>(g, amc, bkc, 1, s);
    case 2: return gemm_dispatch_tile<TC, 128, 128, 16, 2, false>(g, amc, bkc, 1, s);
    case 3: return gemm_dispatch_tile<TC, 64, 64, 16, 2, false>(g, amc, bkc, 1, s);
    case 4: return gemm_dispatch_tile<TC, 64, 128, 16, 4, true, 0>(g, amc, bkc, 1, s);
    default: return hipErrorInvalidValue;
  }
}

int gemm_tile_variants(const GemmDesc& g0, int prec, int* ids) {
  GemmDesc g = g0;
  bool amc, bkc;
  if (prec != PREC_FP64 && prec != PREC_FP32) return 0;
  if (g.M <= 0 || g.N <= 0 || !gemm_normalise(g, amc, bkc)) return 0;
  if (gemm_select_splitk(g, amc, bkc) > 1) return 0;
  int n = 0;
  const bool v128 = gemm_vec_ok(g, amc, bkc, 64, 128);
  if (v128) ids[n++] = 0;
  if (gemm_vec_ok(g, amc, bkc, 64, 64)) ids[n++] = 1;
  ids[n++] = 2;
  ids[n++] = 3;
  if (v128) ids[n++] = 4;
  if (v128) ids[n++] = 5;
  return n;
}

hipError_t launch_gemm_variant(const GemmDesc& g0, int prec, int v, hipStream_t s) {
  GemmDesc g = g0;
  bool amc, bkc;
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  if (!gemm_normalise(g, amc, bkc)) return hipErrorInvalidValue;
  if (((v == 0 || v == 4 || v == 5) && !gemm_vec_ok(g, amc, bkc, 64, 128)) ||
      (v == 1 && !gemm_vec_ok(g, amc, bkc, 64, 64)))
    return hipErrorInvalidValue;
  if (prec == PREC_FP32) return gemm_variant<float>(g, amc, bkc, v, s);
  if (prec == PREC_FP64) return gemm_variant<double>(g, amc, bkc, v, s);
  return hipErrorInvalidValue;
}

hipError_t launch_gemm(const GemmDesc& g0, int prec, hipStream_t s) {
  GemmDesc g = g0;
  if (g.M <= 0 || g.N <= 0) return hipSuccess;
  bool amc, bkc;
  if (!gemm_normalise(g, amc, bkc)) return hipErrorInvalidValue;
  if (prec == PREC_BF16X3) {
    if (g.Kc < g.K && (g.Kc % 32) != 0) return hipErrorInvalidValue;
    const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
    if (t128 >= 192) return gemm_bf16x3_tile<128, 128>(g, amc, bkc, 1, s);
    const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
    int splitk = 1;
    if (g.W != nullptr && t64 < 512) {
      const int ktiles = (g.K + 31) / 32;
      splitk = (int)std::min<long>(8, std::max<long>(1, 1024 / std::max<long>(t64, 1)));
      splitk = std::min(splitk, std::max(1, ktiles / 8));
    }
    return gemm_bf16x3_tile<64, 64>(g, amc, bkc, splitk, s);
  }
  if (prec == PREC_FP32) return gemm_select<float>(g, amc, bkc, s);
  if (prec == PREC_FP64) return gemm_select<double>(g, amc, bkc, s);
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------
// MPO mixing kernels (bandwidth bound)
// ---------------------------------------------------------------------------------

struct Mat8 {
  double v[64];
};

template <int MAXD>
__global__ void mix_mid_kernel(int D1, int d, int D2, Mat8 M, const double* __restrict__ in,
                               double* __restrict__ out, double beta) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (a >= D1) return;
  const size_t base = (size_t)a + (size_t)D1 * d * b;
  double x[MAXD];
#pragma unroll
  for (int s = 0; s < MAXD; ++s)
    if (s < d) x[s] = in[base + (size_t)D1 * s];
  for (int u = 0; u < d; ++u) {
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < MAXD; ++s)
      if (s < d) acc += M.v[u + d * s] * x[s];
    double* p = out + base + (size_t)D1 * u;
    *p = (beta == 0.0) ? acc : acc + beta * *p;
  }
}

hipError_t launch_mix_mid(int D1, int d, int D2, const double* M, const double* in, double* out,
                          double beta, hipStream_t s) {
  if (d > 8) return hipErrorInvalidValue;
  Mat8 m{};
  for (int i = 0; i < d * d; ++i) m.v[i] = M[i];
  dim3 block(256);
  dim3 grid((D1 + 255) / 256, D2);
  hipLaunchKernelGGL((mix_mid_kernel<8>), grid, block, 0, s, D1, d, D2, m, in, out, beta);
  return hipGetLastError();
}

// X0 != nullptr: level i = 0 of T1 is read from X0 (x[b, s, bb] at b + D(s + d bb)) instead;
// Y != nullptr: channel j = w-1 is written to Y (y[b, u, bb] at b + D(u + d bb)) instead of T2.
template <int MAXIN>
__global__ void __launch_bounds__(256) wmix_kernel(int D, int w, int d, const double* __restrict__ Wm,
                                                   const double* __restrict__ T1, double* __restrict__ T2,
                                                   const double* __restrict__ X0, double* __restrict__ Y) {
  __shared__ double Ws[MAXIN * MAXIN];
  const int wd = w * d;
  for (int t = threadIdx.x; t < wd * wd; t += blockDim.x) Ws[t] = Wm[t];
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int bb = blockIdx.y;
  if (b >= D) return;
  double x[MAXIN];
  const double* src = T1 + b + (size_t)D * wd * bb;
  const double* src0 = X0 ? X0 + b + (size_t)D * d * bb : nullptr;
#pragma unroll
  for (int q = 0; q < MAXIN; ++q) {
    if (q < wd) {
      const int i = q % w;
      x[q] = (src0 && i == 0) ? src0[(size_t)D * (q / w)] : src[(size_t)D * q];
    }
  }
  for (int j = 0; j < w; ++j) {
    double* dst = (Y && j == w - 1) ? Y + b + (size_t)D * d * bb : T2 + b + (size_t)D * d * (bb + (size_t)D * j);
    for (int u = 0; u < d; ++u) {
      const int o = u + d * j;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < MAXIN; ++q)
        if (q < wd) acc += Ws[o + wd * q] * x[q];
      dst[(size_t)D * u] = acc;
    }
  }
}

// The same mix on FP64 MFMA tiles (w d <= 32, D % 32 == 0): per bb, OUT (o x b) = W (o x q) .
// IN (q x b) as 16 x 16 x 4 products.  A wave owns 32 consecutive b of one bb (lane li: b0 + 2 li
// and b0 + 2 li + 1, two column tiles), all o (two 16-row tiles) and all q (eight k-steps of 4),
// so every lane moves 16-byte pairs: the W coefficients sit in registers and nothing goes
// through LDS.
__global__ void __launch_bounds__(256) wmix_mfma_kernel(int D, int w, int d, const double* __restrict__ Wm,
                                                        const double* __restrict__ T1, double* __restrict__ T2,
                                                        const double* __restrict__ X0, double* __restrict__ Y) {
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef double v2 __attribute__((ext_vector_type(2)));
  const int wd = w * d;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int bb = blockIdx.y;
  const int b0 = (blockIdx.x * 4 + wave) * 32;
  if (b0 >= D) return;
  // A fragments: lane holds W[o = 16 t + li][q = 4 s + lk] (zero padded to 32 x 32)
  double wa[2][8];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int o = 16 * t + li, q = 4 * s + lk;
      wa[t][s] = (o < wd && q < wd) ? Wm[o + wd * q] : 0.0;
    }
  const int b = b0 + 2 * li;
  v2 xin[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int q = 4 * s + lk;
    xin[s] = v2{0.0, 0.0};
    if (q < wd) {
      const int i = q % w;
      xin[s] = (X0 && i == 0) ? *(const v2*)&X0[b + (size_t)D * (q / w + (size_t)d * bb)]
                              : *(const v2*)&T1[b + (size_t)D * (q + (size_t)wd * bb)];
    }
  }
  v4 acc[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t) acc[t][0] = acc[t][1] = v4{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      acc[t][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[t][s], xin[s][0], acc[t][0], 0, 0, 0);
      acc[t][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(wa[t][s], xin[s][1], acc[t][1], 0, 0, 0);
    }
  // accumulator register r of lane l: o = 16 t + lk + 4 r, b = b0 + 2 li (+1 for the odd tile)
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int o = 16 * t + lk + 4 * r;
      if (o < wd) {
        const int u = o % d, j = o / d;
        double* dst = (Y && j == w - 1) ? Y + b + (size_t)D * (u + (size_t)d * bb)
                                        : T2 + b + (size_t)D * (u + (size_t)d * (bb + (size_t)D * j));
        *(v2*)dst = v2{acc[t][0][r], acc[t][1][r]};
      }
    }
}

hipError_t launch_wmix(int D, int w, int d, const double* Wm_dev, const double* T1, double* T2, const double* X0,
                       double* Y, hipStream_t s) {
  const int wd = w * d;
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  if (wd <= 32 && D % 32 == 0 && al16(T1) && al16(T2) && (!X0 || al16(X0)) && (!Y || al16(Y))) {
    hipLaunchKernelGGL(wmix_mfma_kernel, dim3(D / 128 + (D % 128 ? 1 : 0), D), dim3(256), 0, s, D, w, d, Wm_dev,
                       T1, T2, X0, Y);
    return hipGetLastError();
  }
  dim3 block(256);
  dim3 grid((D + 255) / 256, D);
  if (wd <= 16)
    hipLaunchKernelGGL((wmix_kernel<16>), grid, block, 0, s, D, w, d, Wm_dev, T1, T2, X0, Y);
  else if (wd <= 32)
    hipLaunchKernelGGL((wmix_kernel<32>), grid, block, 0, s, D, w, d, Wm_dev, T1, T2, X0, Y);
  else if (wd <= 64)
    hipLaunchKernelGGL((wmix_kernel<64>), grid, block, 0, s, D, w, d, Wm_dev, T1, T2, X0, Y);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// Energy gradient sandwich: for one bb per workgroup,
//   Gp[o * D + bb] = sum_b T1[b, i, s, bb] * T3[b, u, bb, j],  o = q + wd r, q = i + w s, r = u + d j
// with T1 at b + D(i + w(s + d bb)) and T3 at b + D(u + d(bb + D j)); wd <= 32.
__global__ void __launch_bounds__(256) gsand_kernel(int D, int w, int d, const double* __restrict__ T1,
                                                    const double* __restrict__ T3, double* __restrict__ Gp) {
  __shared__ double S1[64][33];
  __shared__ double S3[64][33];
  const int wd = w * d;
  const int bb = blockIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b0 = 0; b0 < D; b0 += 64) {
    const int nbk = min(64, D - b0);
    for (int t = threadIdx.x; t < nbk * wd; t += blockDim.x) {
      const int bl = t % nbk, q = t / nbk;
      S1[bl][q] = T1[(size_t)(b0 + bl) + (size_t)D * q + (size_t)D * wd * bb];
      const int u = q % d, j = q / d;  // q doubles as r = u + d j here
      S3[bl][q] = T3[(size_t)(b0 + bl) + (size_t)D * u + (size_t)D * d * bb + (size_t)D * D * d * j];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = threadIdx.x + 256 * e;
      if (o < wd * wd) {
        const int q = o % wd, r = o / wd;
        double sacc = 0.0;
        for (int bl = 0; bl < nbk; ++bl) sacc += S1[bl][q] * S3[bl][r];
        acc[e] += sacc;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int o = threadIdx.x + 256 * e;
    if (o < wd * wd) Gp[(size_t)o * D + bb] = acc[e];
  }
}

hipError_t launch_gsand(int D, int w, int d, const double* T1, const double* T3, double* Gp, hipStream_t s) {
  if (w * d > 32) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gsand_kernel, dim3(D), dim3(256), 0, s, D, w, d, T1, T3, Gp);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// FP64 reductions: per-block partial sums, summed on the host in block order.
// ---------------------------------------------------------------------------------

static constexpr int RED_THREADS = 256;
static constexpr int RED_MAXBLK = 1024;

int dot_blocks(size_t n) {
  // about four elements per thread, at most RED_MAXBLK workgroups (sixteen elements per
  // thread for the largest vectors): small vectors still spread over many CUs
  size_t nb = (n + (size_t)RED_THREADS * 4 - 1) / ((size_t)RED_THREADS * 4);
  if (nb < 1) nb = 1;
  if (nb > (size_t)RED_MAXBLK) nb = RED_MAXBLK;
  return (int)nb;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int NV>
__device__ inline void block_sum_store(double (&acc)[NV], int nvalid, double* part, int stride,
                                       int blk) {
  __shared__ double red[NV][RED_THREADS / 64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    const double v = wave_sum(acc[t]);
    if (lane == 0) red[t][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV && (int)threadIdx.x < nvalid) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < RED_THREADS / 64; ++w) s += red[threadIdx.x][w];
    part[(size_t)threadIdx.x * stride + blk] = s;
  }
  __syncthreads();
}

// part[v nb + b] = block b's share of X_v . y: vectors in chunks of 32, so y is read once
// per 32 vectors; each element's loads are all issued before its products (as in the fused
// sweeps below)
__global__ void __launch_bounds__(RED_THREADS) multidot_kernel(size_t n, int nv, const double* __restrict__ X,
                                                               size_t ldx, const double* __restrict__ y,
                                                               double* __restrict__ part) {
  constexpr int CH = 32;
  const int nb = gridDim.x;
  for (int vc = 0; vc < nv; vc += CH) {
    const int nc = min(CH, nv - vc);
    const double* Xc = X + (size_t)vc * ldx;
    double acc[CH];
#pragma unroll
    for (int t = 0; t < CH; ++t) acc[t] = 0.0;
    for (size_t i = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (size_t)nb * RED_THREADS) {
      double xv[CH];
#pragma unroll
      for (int t = 0; t < CH; ++t) xv[t] = (t < nc) ? Xc[(size_t)t * ldx + i] : 0.0;
      const double yi = y[i];
#pragma unroll
      for (int t = 0; t < CH; ++t) acc[t] += xv[t] * yi;
    }
    block_sum_store<CH>(acc, nc, part + (size_t)vc * nb, nb, blockIdx.x);
  }
}

hipError_t launch_multidot(size_t n, int nv, const double* X, size_t ldx, const double* y,
                           double* part, hipStream_t s) {
  const int nb = dot_blocks(n);
  hipLaunchKernelGGL(multidot_kernel, dim3(nb), dim3(RED_THREADS), 0, s, n, nv, X, ldx, y, part);
  return hipGetLastError();
}

static dim3 ew_grid(size_t n);

// ---------------------------------------------------------------------------------
// MINRES on the Jacobi-Davidson correction equation P (H - theta) P t = -r (P = I - x x^T),
// with every scalar of the short recurrence kept on the device (S, see imps_internal.h), so
// that a step needs no host round trip.
// ---------------------------------------------------------------------------------

// part[b] = block b's share of a . y, part[nb + b] = block b's share of b . y
__global__ void __launch_bounds__(RED_THREADS) dot2_kernel(size_t n, const double* __restrict__ a,
                                                           const double* __restrict__ b,
                                                           const double* __restrict__ y, double* __restrict__ part) {
  const int nb = gridDim.x;
  double acc[2] = {0.0, 0.0};
  for (size_t i = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (size_t)nb * RED_THREADS) {
    const double yi = y[i];
    acc[0] += a[i] * yi;
    acc[1] += b[i] * yi;
  }
  block_sum_store<2>(acc, 2, part, nb, blockIdx.x);
}

hipError_t launch_dot2(size_t n, const double* a, const double* b, const double* y, double* part, hipStream_t s) {
  hipLaunchKernelGGL(dot2_kernel, dim3(dot_blocks(n)), dim3(RED_THREADS), 0, s, n, a, b, y, part);
  return hipGetLastError();
}

// z = y - S[GAMMA] x - S[DELTA] v - S[BETA] v_old (the new Lanczos vector of the projected,
// shifted operator before normalisation), part <- block shares of z . z
__global__ void __launch_bounds__(RED_THREADS) minres_z_kernel(size_t n, const double* __restrict__ y,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ v,
                                                               const double* __restrict__ vold,
                                                               const double* __restrict__ S, double* __restrict__ z,
                                                               double* __restrict__ part) {
  const int nb = gridDim.x;
  const double g = S[MR_GAMMA], dl = S[MR_DELTA], bo = S[MR_BETA];
  double acc[1] = {0.0};
  for (size_t i = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (size_t)nb * RED_THREADS) {
    const double zi = y[i] - g * x[i] - dl * v[i] - bo * vold[i];
    z[i] = zi;
    acc[0] += zi * zi;
  }
  block_sum_store<1>(acc, 1, part, nb, blockIdx.x);
}

hipError_t launch_minres_z(size_t n, const double* y, const double* x, const double* v, const double* vold,
                           const double* S, double* z, double* part, hipStream_t s) {
  hipLaunchKernelGGL(minres_z_kernel, dim3(dot_blocks(n)), dim3(RED_THREADS), 0, s, n, y, x, v, vold, S, z, part);
  return hipGetLastError();
}

// One thread: the Givens step of MINRES (Paige-Saunders) from alpha = S[DELTA] - theta and
// beta = sqrt(S[ZZ]); leaves the coefficients of the vector update in S
__global__ void minres_scalar_kernel(double* S, double theta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double alpha = S[MR_DELTA] - theta;
  const double bn = sqrt(fmax(S[MR_ZZ], 0.0));
  const double bo = S[MR_BETA], co = S[MR_COLD], c = S[MR_C], so = S[MR_SOLD], sv = S[MR_S], eta = S[MR_ETA];
  const double rho0 = c * alpha - co * sv * bo;
  const double rho1 = hypot(rho0, bn);
  const double rho2 = sv * alpha + co * c * bo;
  const double rho3 = so * bo;
  if (!(rho1 > 0.0) || S[MR_DONE] != 0.0) {
    S[MR_KV] = S[MR_KWOLD] = S[MR_KW] = S[MR_KT] = S[MR_INVB] = 0.0;
    S[MR_DONE] = 1.0;
    return;
  }
  const double cn = rho0 / rho1, sn = bn / rho1;
  S[MR_KV] = 1.0 / rho1;
  S[MR_KWOLD] = -rho3 / rho1;
  S[MR_KW] = -rho2 / rho1;
  S[MR_KT] = cn * eta;
  S[MR_ETA] = -sn * eta;
  S[MR_COLD] = c;
  S[MR_C] = cn;
  S[MR_SOLD] = sv;
  S[MR_S] = sn;
  S[MR_BETA] = bn;
  S[MR_INVB] = bn > 0.0 ? 1.0 / bn : 0.0;
  if (!(bn > 0.0)) S[MR_DONE] = 1.0;  // invariant subspace: t is exact after this step
}

hipError_t launch_minres_scalar(double* S, double theta, hipStream_t s) {
  hipLaunchKernelGGL(minres_scalar_kernel, dim3(1), dim3(64), 0, s, S, theta);
  return hipGetLastError();
}

// w_new = K_V v + K_WOLD w_old + K_W w (into w_old), t += K_T w_new, z *= INV_BETA (the next v)
__global__ void minres_update_kernel(size_t n, const double* __restrict__ v, double* __restrict__ wold,
                                     const double* __restrict__ w, double* __restrict__ t, double* __restrict__ z,
                                     const double* __restrict__ S) {
  const double kv = S[MR_KV], kwo = S[MR_KWOLD], kw = S[MR_KW], kt = S[MR_KT], ib = S[MR_INVB];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double wn = kv * v[i] + kwo * wold[i] + kw * w[i];
    wold[i] = wn;
    t[i] += kt * wn;
    z[i] *= ib;
  }
}

hipError_t launch_minres_update(size_t n, const double* v, double* wold, const double* w, double* t, double* z,
                                const double* S, hipStream_t s) {
  hipLaunchKernelGGL(minres_update_kernel, ew_grid(n), dim3(256), 0, s, n, v, wold, w, t, z, S);
  return hipGetLastError();
}

// S <- the initial MINRES state for a right-hand side of norm beta1
__global__ void minres_init_kernel(double* S, double beta1) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int k = 0; k < MR_NS; ++k) S[k] = 0.0;
  S[MR_COLD] = 1.0;
  S[MR_C] = 1.0;
  S[MR_ETA] = beta1;
}

hipError_t launch_minres_init(double* S, double beta1, hipStream_t s) {
  hipLaunchKernelGGL(minres_init_kernel, dim3(1), dim3(64), 0, s, S, beta1);
  return hipGetLastError();
}

// Fused Gram-Schmidt sweeps over w (nv <= 32 basis vectors, coefficients on the device):
//   DOTS: w -= X coef, then part <- block shares of X_v . w (the next pass's coefficients)
//   NORM: w -= X coef, then part <- block shares of w . w
// Each basis vector is read once per sweep: the subtraction and the following reduction
// share the loads.
template <bool DOTS>
__global__ void __launch_bounds__(RED_THREADS) axpy_reduce_kernel(size_t n, int nv, const double* __restrict__ X,
                                                                  size_t ldx, const double* __restrict__ coef,
                                                                  double* __restrict__ w, double* __restrict__ part) {
  constexpr int MAXV = 32;
  constexpr int NA = DOTS ? MAXV : 1;
  __shared__ double c[MAXV];
  if (threadIdx.x < MAXV) c[threadIdx.x] = (threadIdx.x < (unsigned)nv) ? coef[threadIdx.x] : 0.0;
  __syncthreads();
  const int nb = gridDim.x;
  double acc[NA];
#pragma unroll
  for (int t = 0; t < NA; ++t) acc[t] = 0.0;
  for (size_t i = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (size_t)nb * RED_THREADS) {
    double xv[MAXV];
    double wi = w[i];
#pragma unroll
    for (int v = 0; v < MAXV; ++v) {
      xv[v] = (v < nv) ? X[(size_t)v * ldx + i] : 0.0;
      wi -= c[v] * xv[v];
    }
    w[i] = wi;
    if (DOTS) {
#pragma unroll
      for (int v = 0; v < NA; ++v) acc[v] += xv[v] * wi;
    } else {
      acc[0] += wi * wi;
    }
  }
  block_sum_store<NA>(acc, DOTS ? nv : 1, part, nb, blockIdx.x);
}

hipError_t launch_axpy_reduce(size_t n, int nv, const double* X, size_t ldx, const double* coef, double* w,
                              double* part, bool dots, hipStream_t s) {
  if (nv > 32 || nv < 1) return hipErrorInvalidValue;
  const int nb = dot_blocks(n);
  if (dots)
    hipLaunchKernelGGL(axpy_reduce_kernel<true>, dim3(nb), dim3(RED_THREADS), 0, s, n, nv, X, ldx, coef, w, part);
  else
    hipLaunchKernelGGL(axpy_reduce_kernel<false>, dim3(nb), dim3(RED_THREADS), 0, s, n, nv, X, ldx, coef, w, part);
  return hipGetLastError();
}

struct Coef64 {
  double c[64];
};

__global__ void multiaxpy_kernel(size_t n, int nv, const double* __restrict__ X, size_t ldx, Coef64 c,
                                 double alpha, double* __restrict__ y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double acc = (alpha == 0.0) ? 0.0 : alpha * y[i];
    for (int v = 0; v < nv; ++v) acc += c.c[v] * X[(size_t)v * ldx + i];
    y[i] = acc;
  }
}

static dim3 ew_grid(size_t n) {
  size_t nb = (n + 255) / 256;
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  return dim3((unsigned)nb);
}

hipError_t launch_multiaxpy(size_t n, int nv, const double* X, size_t ldx, const double* coef,
                            double alpha, double* y, hipStream_t s) {
  if (nv > 64) return hipErrorInvalidValue;
  Coef64 c{};
  for (int v = 0; v < nv; ++v) c.c[v] = coef[v];
  hipLaunchKernelGGL(multiaxpy_kernel, ew_grid(n), dim3(256), 0, s, n, nv, X, ldx, c, alpha, y);
  return hipGetLastError();
}

// out[v] = sum_b part[v nb + b]: one workgroup per v, strided partial sums, then an LDS tree
// (fixed order, so repeated calls give identical results)
__global__ void __launch_bounds__(256) partials_reduce_kernel(int nv, int nb, const double* __restrict__ part,
                                                             double* __restrict__ out) {
  __shared__ double red[256];
  const int v = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += part[(size_t)v * nb + b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[v] = red[0];
}

hipError_t launch_partials_reduce(int nv, int nb, const double* part, double* out, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  hipLaunchKernelGGL(partials_reduce_kernel, dim3(nv), dim3(256), 0, s, nv, nb, part, out);
  return hipGetLastError();
}

__global__ void multiaxpy_dev_kernel(size_t n, int nv, const double* __restrict__ X, size_t ldx,
                                     const double* __restrict__ coef, double scale, double alpha,
                                     double* __restrict__ y) {
  __shared__ double c[64];
  if (threadIdx.x < nv) c[threadIdx.x] = scale * coef[threadIdx.x];
  __syncthreads();
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double acc = (alpha == 0.0) ? 0.0 : alpha * y[i];
    for (int v = 0; v < nv; ++v) acc += c[v] * X[(size_t)v * ldx + i];
    y[i] = acc;
  }
}

__global__ void scale_inv_sqrt_kernel(size_t n, double* x, const double* sq) {
  const double q = *sq;
  if (!(q > 0.0)) return;
  const double a = 1.0 / sqrt(q);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

__global__ void scale_kernel(size_t n, double a, double* x) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] *= a;
}

hipError_t launch_multiaxpy_dev(size_t n, int nv, const double* X, size_t ldx, const double* coef,
                                double scale, double alpha, double* y, hipStream_t s) {
  if (nv > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(multiaxpy_dev_kernel, ew_grid(n), dim3(256), 0, s, n, nv, X, ldx, coef, scale, alpha, y);
  return hipGetLastError();
}

hipError_t launch_scale_inv_sqrt(size_t n, double* x, const double* sq, hipStream_t s) {
  hipLaunchKernelGGL(scale_inv_sqrt_kernel, ew_grid(n), dim3(256), 0, s, n, x, sq);
  return hipGetLastError();
}

hipError_t launch_scale(size_t n, double a, double* x, hipStream_t s) {
  hipLaunchKernelGGL(scale_kernel, ew_grid(n), dim3(256), 0, s, n, a, x);
  return hipGetLastError();
}

__global__ void axpby_kernel(size_t n, double a, const double* x, double b, double* y) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (b == 0.0) ? a * x[i] : a * x[i] + b * y[i];
}

hipError_t launch_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t s) {
  hipLaunchKernelGGL(axpby_kernel, ew_grid(n), dim3(256), 0, s, n, a, x, b, y);
  return hipGetLastError();
}

__global__ void fill_kernel(size_t n, double v, double* x) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    x[i] = v;
}

hipError_t launch_fill(size_t n, double v, double* x, hipStream_t s) {
  hipLaunchKernelGGL(fill_kernel, ew_grid(n), dim3(256), 0, s, n, v, x);
  return hipGetLastError();
}

__global__ void add_diag_kernel(int n, double v, double* x, long ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[(size_t)i * ld + i] += v;
}

hipError_t launch_add_diag(int n, double v, double* x, long ld, hipStream_t s) {
  hipLaunchKernelGGL(add_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, v, x, ld);
  return hipGetLastError();
}

__global__ void fill2d_kernel(int rows, int cols, double v, double* x, long ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < rows && j < cols) x[i + ld * j] = v;
}

hipError_t launch_fill2d(int rows, int cols, double v, double* x, long ld, hipStream_t s) {
  hipLaunchKernelGGL(fill2d_kernel, dim3((rows + 255) / 256, cols), dim3(256), 0, s, rows, cols, v, x, ld);
  return hipGetLastError();
}

__global__ void set_identity_kernel(int n, double* x, long ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) x[i + ld * j] = (i == j) ? 1.0 : 0.0;
}

hipError_t launch_set_identity(int n, double* x, long ld, hipStream_t s) {
  hipLaunchKernelGGL(set_identity_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, x, ld);
  return hipGetLastError();
}

__global__ void transpose_kernel(int rows, int cols, const double* __restrict__ in, long ldin,
                                 double* __restrict__ out, long ldout) {
  __shared__ double tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + tx, c = c0 + k;
    if (r < rows && c < cols) tile[k][tx] = in[r + ldin * c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + tx, r = r0 + k;  // out[c + ldout * r] = in[r, c]
    if (r < rows && c < cols) out[c + ldout * r] = tile[tx][k];
  }
}

hipError_t launch_transpose(int rows, int cols, const double* in, long ldin, double* out,
                            long ldout, hipStream_t s) {
  dim3 grid((rows + 31) / 32, (cols + 31) / 32);
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, s, rows, cols, in, ldin, out, ldout);
  return hipGetLastError();
}

// sum_{a,b} l[a + ldl*b] * v[b + ldv*a], one partial per 32 x 32 tile
__global__ void trace_prod_kernel(int n, const double* __restrict__ l, long ldl, const double* __restrict__ v,
                                  long ldv, double* __restrict__ part) {
  __shared__ double vt[32][33];
  __shared__ double red[4];
  const int a0 = blockIdx.x * 32, b0 = blockIdx.y * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
  // v[b + ldv*a] for b in tile rows (contiguous), a in tile columns
  for (int k = ty; k < 32; k += 8) {
    const int b = b0 + tx, a = a0 + k;
    vt[k][tx] = (a < n && b < n) ? v[b + ldv * a] : 0.0;  // vt[a-a0][b-b0]
  }
  __syncthreads();
  double acc = 0.0;
  for (int k = ty; k < 32; k += 8) {
    const int a = a0 + tx, b = b0 + k;
    if (a < n && b < n) acc += l[a + ldl * b] * vt[tx][k];
  }
  const int t = ty * 32 + tx;
  acc = wave_sum(acc);
  if ((t & 63) == 0) red[t >> 6] = acc;
  __syncthreads();
  if (t == 0) part[blockIdx.x + gridDim.x * blockIdx.y] = red[0] + red[1] + red[2] + red[3];
}

hipError_t launch_trace_prod_parts(int n, const double* l, long ldl, const double* v, long ldv, double* part,
                                   int* nparts, hipStream_t s) {
  dim3 grid((n + 31) / 32, (n + 31) / 32);
  *nparts = (int)(grid.x * grid.y);
  hipLaunchKernelGGL(trace_prod_kernel, grid, dim3(32, 8), 0, s, n, l, ldl, v, ldv, part);
  return hipGetLastError();
}

// One workgroup: c = sum_b part[b] (the fixed order of partials_reduce_kernel), then
// x[i + ld*i] += factor * c for i < n; c is stored to *out when out != nullptr.
__global__ void __launch_bounds__(256) reduce_add_diag_kernel(int np, const double* __restrict__ part, double factor,
                                                             double* __restrict__ x, long ld, int n,
                                                             double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < np; b += 256) s += part[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double c = red[0];
  if (out != nullptr && threadIdx.x == 0) *out = c;
  for (int i = threadIdx.x; i < n; i += 256) x[(size_t)i * ld + i] += factor * c;
}

hipError_t launch_reduce_add_diag(int np, const double* part, double factor, double* x, long ld, int n, double* out,
                                  hipStream_t s) {
  hipLaunchKernelGGL(reduce_add_diag_kernel, dim3(1), dim3(256), 0, s, np, part, factor, x, ld, n, out);
  return hipGetLastError();
}

// c = sum_b part[b] (fixed order; every workgroup sums the partials itself), then
// v[i + ldv*j] += factor * c * fix[i + ldf*j] for i, j < n.  Grid: (ceil(n / 256), n).
__global__ void __launch_bounds__(256) reduce_axpy_mat_kernel(int np, const double* __restrict__ part, double factor,
                                                              const double* __restrict__ fix, long ldf,
                                                              double* __restrict__ v, long ldv, int n) {
  __shared__ double red[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < np; b += 256) s += part[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double f = factor * red[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) v[i + ldv * j] += f * fix[i + ldf * j];
}

hipError_t launch_reduce_axpy_mat(int np, const double* part, double factor, const double* fix, long ldf, double* v,
                                  long ldv, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(reduce_axpy_mat_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, np, part, factor, fix, ldf,
                     v, ldv, n);
  return hipGetLastError();
}

__device__ inline double sgn_nonneg(double x) { return x < 0.0 ? -1.0 : 1.0; }

__global__ void qr_signs_kernel(int n, const double* R, long ldr, double* sg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sg[i] = sgn_nonneg(R[i + ldr * i]);
}

__global__ void qr_signfix_q_kernel(int m, int n, double* Q, long ldq, const double* sg) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (r < m && i < n) Q[r + ldq * i] *= sg[i];
}

__global__ void qr_signfix_r_kernel(int n, double* R, long ldr, const double* sg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // row
  const int j = blockIdx.y;                             // column
  if (i < n && i <= j) R[i + ldr * j] *= sg[i];
}

hipError_t launch_qr_signfix(int m, int n, double* Q, long ldq, double* R, long ldr, double* sbuf,
                             hipStream_t s) {
  hipLaunchKernelGGL(qr_signs_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, R, ldr, sbuf);
  if (Q) hipLaunchKernelGGL(qr_signfix_q_kernel, dim3((m + 255) / 256, n), dim3(256), 0, s, m, n, Q, ldq, sbuf);
  hipLaunchKernelGGL(qr_signfix_r_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, R, ldr, sbuf);
  return hipGetLastError();
}

// Cholesky of one diagonal block (nb <= 64, padded to 64 x 64 with an identity block) and the
// inverse of its factor, by four wavefronts in lockstep over the 64 right-looking steps (one
// barrier per step).  Step k: the wave owning column k publishes it, not yet scaled, as
// c = A^(k)[:, k]; every wave reads the pivot c_k from LDS.
//  - factor waves (lane i = row i; columns 0-31 and 32-63): L[i][k] = c_i / sqrt(c_k) and the
//    rank-1 update A[i][j] -= c_i c_j / c_k for the wave's columns j > k;
//  - inverse waves (lane j = column j of X = L^{-1}; rows 0-31 and 32-63): X[k][j] =
//    x_j[k] / L[k][k] and x_j[i] -= L[i][k] X[k][j] for the wave's rows i > k.  The rows-32-63
//    wave applies the steps k < 32 one step late, with X[k][j] / L[k][k] handed over in LDS.
// The column buffer is triple-buffered (the late wave still reads column k-1 while column k+1
// is published) and the hand-over buffer double-buffered.  L goes back to A (lower triangle);
// U = L^{-T} (64 x 64, zero below the diagonal and beyond nb) goes to Ut, so that the panel
// solve X = P L^{-T} is the product P U.
template <int J0>
__device__ inline void chol_fact_wave(int lane, double (*c)[64], double (*Ls)[65], int* fail) {
  double v[32];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < 32; ++j) v[j] = Ls[J0 + j][lane];
#pragma clang loop unroll(full)
  for (int k = 0; k < 64; ++k) {
    double* ck = c[k % 3];
    const bool own = (k >= J0 && k < J0 + 32);
    if (own) ck[lane] = v[k - J0 < 0 ? 0 : (k - J0 > 31 ? 31 : k - J0)];
    __syncthreads();
    if (!own && k >= J0 + 31) continue;  // no column of this wave at or right of k
    double p = ck[k];
    if (!(p > 0.0)) {
      bad = true;
      p = 1.0;
    }
    const double inv = rsqrt(p);
    const double ci = ck[lane];
    const double t = (lane > k) ? ci * (inv * inv) : 0.0;
    if (own) Ls[k][lane] = (lane > k) ? ci * inv : (lane == k ? p * inv : 0.0);
#pragma unroll
    for (int j = 0; j < 32; ++j)
      if (J0 + j > k) v[j] -= t * ck[J0 + j];
  }
  if (bad && lane == 0) *fail = 1;
}

template <int I0>
__device__ inline void chol_inv_wave(int lane, double (*c)[64], double (*scb)[64], double (*Us)[65]) {
  double v[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) v[i] = (lane == I0 + i) ? 1.0 : 0.0;
#pragma clang loop unroll(full)
  for (int k = 0; k < 64; ++k) {
    __syncthreads();
    if (I0 == 32 && k >= 1 && k - 1 < 32) {
      const double* cp = c[(k - 1) % 3];
      const double sc = scb[(k - 1) & 1][lane];
#pragma unroll
      for (int i = 0; i < 32; ++i) v[i] -= cp[32 + i] * sc;
    }
    const bool own = (k >= I0 && k < I0 + 32);
    if (!own) continue;
    const double* ck = c[k % 3];
    double p = ck[k];
    if (!(p > 0.0)) p = 1.0;
    const double inv = rsqrt(p);
    const double y = v[k - I0 < 0 ? 0 : (k - I0 > 31 ? 31 : k - I0)] * inv;
    Us[k][lane] = y;
    const double sc = y * inv;
    if (I0 == 0) scb[k & 1][lane] = sc;
#pragma unroll
    for (int i = 0; i < 32; ++i)
      if (I0 + i > k) v[i] -= ck[I0 + i] * sc;
  }
}

__global__ void __launch_bounds__(256) chol_diag_kernel(int nb, double* A, long lda, double* Ut, int* fail) {
  __shared__ double c[3][64];
  __shared__ double scb[2][64];
  __shared__ double Ls[64][65];  // Ls[k][i] = block entry (i, k): input, then L
  __shared__ double Us[64][65];  // Us[k][j] = X[k][j] = U[j][k]
  const int lane = threadIdx.x & 63;
  for (int t = threadIdx.x; t < 64 * 64; t += 256) {
    const int i = t & 63, j = t >> 6;
    double x = (i == j) ? 1.0 : 0.0;
    if (i < nb && j < nb && i >= j) x = A[i + lda * j];
    Ls[j][i] = x;
  }
  __syncthreads();
  // wave-uniform role: each wave runs its own straight-line loop; all meet at the same 64
  // barriers
  const int role = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (role == 0)
    chol_fact_wave<0>(lane, c, Ls, fail);
  else if (role == 1)
    chol_fact_wave<32>(lane, c, Ls, fail);
  else if (role == 2)
    chol_inv_wave<0>(lane, c, scb, Us);
  else
    chol_inv_wave<32>(lane, c, scb, Us);
  __syncthreads();
  for (int t = threadIdx.x; t < 64 * 64; t += 256) {
    const int i = t & 63, j = t >> 6;
    if (i < nb && j < nb && i >= j) A[i + lda * j] = Ls[j][i];
    Ut[t] = (i < nb && j < nb) ? Us[j][i] : 0.0;
  }
}

hipError_t launch_chol_diag(int nb, double* A, long lda, double* Ut, int* fail, hipStream_t s) {
  if (nb > 64 || nb < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(256), 0, s, nb, A, lda, Ut, fail);
  return hipGetLastError();
}

// G[i][i] += factor * trace(G) for an n x n matrix (one workgroup: the trace, then the shift)
__global__ void __launch_bounds__(256) add_diag_trace_kernel(int n, double factor, double* G, long ld) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += G[i + ld * i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double shift = factor * red[0];
  for (int i = threadIdx.x; i < n; i += 256) G[i + ld * i] += shift;
}

hipError_t launch_add_diag_trace(int n, double factor, double* G, long ld, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(add_diag_trace_kernel, dim3(1), dim3(256), 0, s, n, factor, G, ld);
  return hipGetLastError();
}

__global__ void mirror_lower_kernel(int n, double* A, long ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n && i > j) A[j + ld * i] = A[i + ld * j];
}

hipError_t launch_mirror_lower(int n, double* A, long ld, hipStream_t s) {
  hipLaunchKernelGGL(mirror_lower_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, A, ld);
  return hipGetLastError();
}

__global__ void fill_int_kernel(int n, int v, int* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

hipError_t launch_fill_int(int n, int v, int* x, hipStream_t s) {
  hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, v, x);
  return hipGetLastError();
}

__global__ void tri_keep_kernel(int n, double* A, long ld, bool keep_upper) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n && ((keep_upper && i > j) || (!keep_upper && i < j))) A[i + ld * j] = 0.0;
}

hipError_t launch_tri_keep(int n, double* A, long ld, bool keep_upper, hipStream_t s) {
  hipLaunchKernelGGL(tri_keep_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, A, ld, keep_upper);
  return hipGetLastError();
}

__global__ void copy_upper_kernel(int n, const double* A, long lda, double* R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < n) R[i + (long)n * j] = (i <= j) ? A[i + lda * j] : 0.0;
}

hipError_t launch_copy_upper(int n, const double* A, long lda, double* R, hipStream_t s) {
  hipLaunchKernelGGL(copy_upper_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, A, lda, R);
  return hipGetLastError();
}

__global__ void copy2d_kernel(int rows, int cols, const double* src, long lds, double* dst, long ldd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < rows && j < cols) dst[i + ldd * j] = src[i + lds * j];
}

hipError_t launch_copy2d(int rows, int cols, const double* src, long lds, double* dst, long ldd,
                         hipStream_t s) {
  hipLaunchKernelGGL(copy2d_kernel, dim3((rows + 255) / 256, cols), dim3(256), 0, s, rows, cols, src, lds, dst, ldd);
  return hipGetLastError();
}

__global__ void fill_random_kernel(size_t n, uint64_t seed, double* x) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

hipError_t launch_fill_random(size_t n, uint64_t seed, double* x, hipStream_t s) {
  hipLaunchKernelGGL(fill_random_kernel, ew_grid(n), dim3(256), 0, s, n, seed, x);
  return hipGetLastError();
}

// ---- complex states on the real kernels: each complex entry z of a (D1, m, D2) tensor becomes
// the real 2x2 block [[Re z, -Im z], [Im z, Re z]] at rows 2a, 2a+1 and columns 2b, 2b+1 of a
// (2 D1, m, 2 D2) tensor.  The map is an algebra homomorphism (products, adjoints as transposes,
// and the positive-diagonal QR/LQ/Cholesky factors all commute with it), so the real algorithms
// run unchanged on the embedded tensors.

// R = scale * embed(Z); Z complex (re, im interleaved), column-major (D1, m, D2)
__global__ void cplx_embed_kernel(int D1, int m, int D2, const double* __restrict__ Z, double scale,
                                  double* __restrict__ R) {
  const size_t n = (size_t)D1 * m * D2;
  const size_t ldr = (size_t)2 * D1;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t a = idx % D1, kb = idx / D1, k = kb % m, b = kb / m;
    const double re = scale * Z[2 * idx], im = scale * Z[2 * idx + 1];
    const size_t c0 = ldr * (k + (size_t)m * (2 * b)), c1 = ldr * (k + (size_t)m * (2 * b + 1));
    R[2 * a + c0] = re;
    R[2 * a + 1 + c0] = im;
    R[2 * a + c1] = -im;
    R[2 * a + 1 + c1] = re;
  }
}

// Z = scale * (the complex tensor nearest to R): each 2x2 block [[p, q], [r, t]] gives
// ((p + t) / 2, (r - q) / 2); with project, R is also overwritten by the embedding of that value
__global__ void cplx_extract_kernel(int D1, int m, int D2, double* __restrict__ R, double scale,
                                    double* __restrict__ Z, bool project) {
  const size_t n = (size_t)D1 * m * D2;
  const size_t ldr = (size_t)2 * D1;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const size_t a = idx % D1, kb = idx / D1, k = kb % m, b = kb / m;
    const size_t c0 = ldr * (k + (size_t)m * (2 * b)), c1 = ldr * (k + (size_t)m * (2 * b + 1));
    const double p = R[2 * a + c0], r = R[2 * a + 1 + c0], q = R[2 * a + c1], t = R[2 * a + 1 + c1];
    const double re = 0.5 * (p + t), im = 0.5 * (r - q);
    if (Z) {
      Z[2 * idx] = scale * re;
      Z[2 * idx + 1] = scale * im;
    }
    if (project) {
      R[2 * a + c0] = re;
      R[2 * a + 1 + c0] = im;
      R[2 * a + c1] = -im;
      R[2 * a + 1 + c1] = re;
    }
  }
}

// y += alpha * i w on embedded tensors whose leading (row) dimension is even: multiplying by i is
// the left product with [[0, -1], [1, 0]] on each row pair
__global__ void cplx_axpy_i_kernel(size_t npairs, double alpha, const double* __restrict__ w, double* __restrict__ y) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < npairs; k += (size_t)gridDim.x * blockDim.x) {
    const double w0 = w[2 * k], w1 = w[2 * k + 1];
    y[2 * k] -= alpha * w1;
    y[2 * k + 1] += alpha * w0;
  }
}

hipError_t launch_cplx_embed(int D1, int m, int D2, const double* Z, double scale, double* R, hipStream_t s) {
  hipLaunchKernelGGL(cplx_embed_kernel, ew_grid((size_t)D1 * m * D2), dim3(256), 0, s, D1, m, D2, Z, scale, R);
  return hipGetLastError();
}

hipError_t launch_cplx_extract(int D1, int m, int D2, double* R, double scale, double* Z, bool project,
                               hipStream_t s) {
  hipLaunchKernelGGL(cplx_extract_kernel, ew_grid((size_t)D1 * m * D2), dim3(256), 0, s, D1, m, D2, R, scale, Z,
                     project);
  return hipGetLastError();
}

hipError_t launch_cplx_axpy_i(size_t n, double alpha, const double* w, double* y, hipStream_t s) {
  hipLaunchKernelGGL(cplx_axpy_i_kernel, ew_grid(n / 2), dim3(256), 0, s, n / 2, alpha, w, y);
  return hipGetLastError();
}

// ---- bond dimensions that differ from bond to bond: the state is held at the largest bond
// dimension with zeros outside each bond's active block

// G[i + ld*i] = 1 for nact <= i < n (the Gram matrix of a matrix whose columns from nact on are
// zero: the Cholesky factor then carries an identity there and the solves leave those columns zero)
__global__ void pad_unit_diag_kernel(int n, int nact, double* G, long ld) {
  const int i = nact + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) G[i + ld * i] = 1.0;
}

hipError_t launch_pad_unit_diag(int n, int nact, double* G, long ld, hipStream_t s) {
  if (nact >= n) return hipSuccess;
  hipLaunchKernelGGL(pad_unit_diag_kernel, dim3((n - nact + 255) / 256), dim3(256), 0, s, n, nact, G, ld);
  return hipGetLastError();
}

// A[i + ld*j] = 0 for i >= ract or j >= cact (rows x cols)
__global__ void mask_block_kernel(int rows, int cols, int ract, int cact, double* A, long ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i < rows && j < cols && (i >= ract || j >= cact)) A[i + ld * j] = 0.0;
}

hipError_t launch_mask_block(int rows, int cols, int ract, int cact, double* A, long ld, hipStream_t s) {
  if (ract >= rows && cact >= cols) return hipSuccess;
  hipLaunchKernelGGL(mask_block_kernel, dim3((rows + 255) / 256, cols), dim3(256), 0, s, rows, cols, ract, cact, A,
                     ld);
  return hipGetLastError();
}

}  // namespace imps
