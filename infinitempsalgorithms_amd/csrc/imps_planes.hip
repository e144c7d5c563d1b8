// Split-BF16 "planes" path of H_AC (effective.jl:90-102) for the mixed-precision bulk
// contractions: every FP64 operand x is stored once as two bf16 planes, x_hi = bf16(x) and
// x_lo = bf16(x - x_hi), laid out K-contiguous (row-major over the contraction index), and the
// GEMMs stream those planes into v_mfma_f32_16x16x32_bf16 as a_hi b_hi + a_hi b_lo + a_lo b_hi
// with FP32 accumulation (the same three-term product as the on-the-fly split GEMM).
//
//   split   : FP64 -> (hi, lo) planes, optionally transposing (GL levels -> [m][k] rows)
//   gemm    : (register-staged tiles, or an LDS-DMA ring for long K on small grids)
//             C(m, n) = alpha sum_k A(m, k) B(k, n) (+ beta C), A(m, k) at m lda + k,
//             B(k, n) at n ldb + k, C(m, n) at m + n ldc in FP32 or FP64
//   wmix    : the MPO mix of H_AC between the two GEMMs, reading GEMM 1's FP32 output and
//             writing GEMM 2's A operand as planes (transposed through LDS so that its rows are
//             contiguous along the contraction index)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "imps_internal.h"

namespace imps {

namespace {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

__device__ inline void split_f32(float xf, __bf16& hi, __bf16& lo) {
  hi = (__bf16)xf;
  lo = (__bf16)(xf - (float)hi);
}

// out[r ldo + c] = split(in[r ldi + c]); cols % 8 == 0; one thread per 8 consecutive c
__global__ void __launch_bounds__(256) split_rows_kernel(int rows, int cols, const double* __restrict__ in,
                                                         long ldi, __bf16* __restrict__ hi, __bf16* __restrict__ lo,
                                                         long ldo) {
  const int cpr = cols / 8;
  const long total = (long)rows * cpr;
  for (long t = blockIdx.x * 256L + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
    const int r = (int)(t / cpr);
    const int c = 8 * (int)(t - (long)r * cpr);
    const double* p = in + (long)r * ldi + c;
    v8bf h, l;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      const v2d v = *(const v2d*)(p + e);
      __bf16 a, b;
      split_f32((float)v[0], a, b);
      h[e] = a;
      l[e] = b;
      split_f32((float)v[1], a, b);
      h[e + 1] = a;
      l[e + 1] = b;
    }
    *(v8bf*)(hi + (long)r * ldo + c) = h;
    *(v8bf*)(lo + (long)r * ldo + c) = l;
  }
}

// out[r ldo + c] = split(in[r + ldi c]) through a 64 x 64 LDS tile; rows % 64 == cols % 64 == 0;
// blockIdx.z selects a level: in += z in_z, out += z out_z
template <typename TI>
__global__ void __launch_bounds__(256) split_trans_kernel(const TI* __restrict__ in, long ldi, long in_z,
                                                          __bf16* __restrict__ hi, __bf16* __restrict__ lo,
                                                          long ldo, long out_z) {
  __shared__ float T[64][65];
  const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
  const int t = threadIdx.x;
  in += blockIdx.z * in_z;
  hi += blockIdx.z * out_z;
  lo += blockIdx.z * out_z;
  // read: lanes along r (contiguous in `in`)
  for (int e = 0; e < 16; ++e) {
    const int c = (t >> 6) + 4 * e, r = t & 63;
    T[c][r] = (float)in[(long)(r0 + r) + ldi * (c0 + c)];
  }
  __syncthreads();
  // write: thread -> row r, 16 consecutive c
  const int r = t >> 2, cq = 16 * (t & 3);
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    v8bf h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      __bf16 a, b;
      split_f32(T[cq + 8 * half + e][r], a, b);
      h[e] = a;
      l[e] = b;
    }
    const long o = (long)(r0 + r) * ldo + c0 + cq + 8 * half;
    *(v8bf*)(hi + o) = h;
    *(v8bf*)(lo + o) = l;
  }
}

struct PlaneGemm {
  int M, N, K;
  const __bf16 *Ah, *Al;
  long lda;
  const __bf16 *Bh, *Bl;
  long ldb;
  void* C;
  long ldc;
  float alpha;
  double beta;
  long zstride;  // split-K: blockIdx.z covers K [z K, (z + 1) K) and writes C + z zstride
};

// Block = 4 waves (2 x 2), tile BM x BN, K stage BK (multiple of 32).  Each thread moves
// 16-byte chunks (8 bf16 along k) of both planes global -> registers -> LDS ([row][k], 8 bf16
// of padding per row: the 16 rows read by one ds_read_b128 fall on disjoint banks); the next
// stage is fetched into registers while the MFMAs consume the current one.  The B fragment is
// the MFMA's row operand, so the 16 lanes of an accumulator row hold 16 consecutive m.
template <int BM, int BN, int BK, typename TO>
__global__ void __launch_bounds__(256, 2) gemm_planes_kernel(PlaneGemm g) {
  constexpr int LDK = BK + 8;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int CH = BK / 8;          // 16-byte chunks per row and stage
  constexpr int RSTEP = 256 / CH;     // rows covered by one pass of the block
  constexpr int AE = BM / RSTEP, BE = BN / RSTEP;
  static_assert(AE * RSTEP == BM && BE * RSTEP == BN && BK % 32 == 0, "tile/thread mismatch");
  __shared__ __attribute__((aligned(16))) __bf16 Ash[BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Asl[BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bsh[BN][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bsl[BN][LDK];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
  const int crow = tid / CH, ck = 8 * (tid % CH);

  // wave-uniform stage bases (bytes) plus one 32-bit offset per chunk (checked on the host)
  const long kz = (long)blockIdx.z * g.K;
  const char* pah = (const char*)(g.Ah + (long)m0 * g.lda + kz);
  const char* pal = (const char*)(g.Al + (long)m0 * g.lda + kz);
  const char* pbh = (const char*)(g.Bh + (long)n0 * g.ldb + kz);
  const char* pbl = (const char*)(g.Bl + (long)n0 * g.ldb + kz);
  unsigned offA[AE], offB[BE];
#pragma unroll
  for (int e = 0; e < AE; ++e) offA[e] = (unsigned)(2 * ((long)(crow + e * RSTEP) * g.lda + ck));
#pragma unroll
  for (int e = 0; e < BE; ++e) offB[e] = (unsigned)(2 * ((long)(crow + e * RSTEP) * g.ldb + ck));

  v8bf rah[AE], ral[AE], rbh[BE], rbl[BE];
  auto load = [&]() {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      rah[e] = *(const v8bf*)(pah + offA[e]);
      ral[e] = *(const v8bf*)(pal + offA[e]);
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      rbh[e] = *(const v8bf*)(pbh + offB[e]);
      rbl[e] = *(const v8bf*)(pbl + offB[e]);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      *(v8bf*)&Ash[crow + e * RSTEP][ck] = rah[e];
      *(v8bf*)&Asl[crow + e * RSTEP][ck] = ral[e];
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      *(v8bf*)&Bsh[crow + e * RSTEP][ck] = rbh[e];
      *(v8bf*)&Bsl[crow + e * RSTEP][ck] = rbl[e];
    }
  };

  v4f acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  const int lr = lane & 15, lk = lane >> 4;
  const int nk = g.K / BK;
  load();
  for (int kt = 0; kt < nk; ++kt) {
    __syncthreads();
    store();
    __syncthreads();
    if (kt + 1 < nk) {
      pah += 2 * BK;
      pal += 2 * BK;
      pbh += 2 * BK;
      pbl += 2 * BK;
      load();
    }
#pragma unroll
    for (int k32 = 0; k32 < BK; k32 += 32) {
      v8bf ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        ah[i] = *(const v8bf*)&Ash[wm + 16 * i + lr][k32 + 8 * lk];
        al[i] = *(const v8bf*)&Asl[wm + 16 * i + lr][k32 + 8 * lk];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        bh[j] = *(const v8bf*)&Bsh[wn + 16 * j + lr][k32 + 8 * lk];
        bl[j] = *(const v8bf*)&Bsl[wn + 16 * j + lr][k32 + 8 * lk];
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
  }
  // accumulator register r of lane l: m = l % 16, n = 4 (l / 16) + r within the 16 x 16 tile.
  // With beta the old values of a row of tiles are loaded together before any store.
  TO* C = (TO*)g.C + blockIdx.z * g.zstride;
  const bool with_beta = sizeof(TO) == 8 && g.beta != 0.0;
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int m = m0 + wm + 16 * i + lr;
    TO* p[TN][4];
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) p[j][r] = C + m + (long)(n0 + wn + 16 * j + 4 * lk + r) * g.ldc;
    if (with_beta) {
      double old[TN][4];
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) old[j][r] = (double)*p[j][r];
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) *p[j][r] = (TO)((double)g.alpha * (double)acc[i][j][r] + g.beta * old[j][r]);
    } else {
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) *p[j][r] = (TO)(g.alpha * acc[i][j][r]);
    }
  }
}

// Long-K variant for grids of about one workgroup per CU: 128 x 128 tile, four waves of 64 x 64,
// v_mfma_f32_32x32x16_bf16, and the operand stages fetched straight into LDS
// (global_load_lds_dwordx4: lane L of a wave writes 16 bytes at the wave-uniform LDS address
// + 16 L) through a ring of NS stages of BK = 32, so NS - 1 stages are in flight while one is
// consumed.  One DMA instruction fills 16 rows x 4 chunks; chunk c of row r is stored in slot
// c ^ swz(r), which keeps the 8- and 16-lane groups of a fragment read on disjoint banks.  The
// stage barrier is a bare s_barrier: __syncthreads() would also wait for every outstanding DMA
// (vmcnt(0)) and empty the ring; each wave instead waits for its own DMAs of the stage about to
// be read, then the barrier makes all of them visible.
__device__ __forceinline__ int ring_swz(int row) {
  return (((row >> 2) & 1) << 1) | (((row >> 1) ^ (row >> 3)) & 1);
}

// s_waitcnt vmcnt(N) leaving the other counters alone (expcnt 7, lgkmcnt 15)
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

template <int NS, typename TO>
__global__ void __launch_bounds__(256, 1) gemm_planes_ring_kernel(PlaneGemm g) {
  constexpr int BM = 128, BN = 128, BK = 32, PL = 128 * BK;
  constexpr int PER_STAGE = 8;  // DMA instructions per wave and stage (2 blocks x 4 planes)
  static_assert(NS >= 2 && NS <= 4, "ring depth");
  __shared__ __attribute__((aligned(1024))) __bf16 ring[NS][4][PL];
  typedef float v16f __attribute__((ext_vector_type(16)));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  // this wave fills blocks q = wave and wave + 4 (rows 16 q .. 16 q + 15) of every plane
  unsigned offA[2], offB[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int r = 16 * (wave + 4 * e) + (lane >> 2);
    const int c = (lane & 3) ^ ring_swz(r);
    offA[e] = (unsigned)(2 * ((long)r * g.lda + 8 * c));
    offB[e] = (unsigned)(2 * ((long)r * g.ldb + 8 * c));
  }
  const char* pah = (const char*)(g.Ah + (long)m0 * g.lda);
  const char* pal = (const char*)(g.Al + (long)m0 * g.lda);
  const char* pbh = (const char*)(g.Bh + (long)n0 * g.ldb);
  const char* pbl = (const char*)(g.Bl + (long)n0 * g.ldb);
  auto issue = [&](int stage, int buf) {
    const long kb = 2L * BK * stage;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int q = 16 * BK * (wave + 4 * e);
      __builtin_amdgcn_global_load_lds((const void*)(pah + kb + offA[e]), (void*)&ring[buf][0][q], 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(pal + kb + offA[e]), (void*)&ring[buf][1][q], 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(pbh + kb + offB[e]), (void*)&ring[buf][2][q], 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(pbl + kb + offB[e]), (void*)&ring[buf][3][q], 16, 0, 0);
    }
  };
  v16f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  // fragment (row, 8-element chunk) offsets of this lane for the two k16 halves of a stage
  const int lr = lane & 31, lk = lane >> 5;
  int fa[2][2], fb[2][2];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int ra = wm + 32 * i + lr, rb = wn + 32 * i + lr;
      const int c = 2 * h + lk;
      fa[h][i] = ra * BK + 8 * (c ^ ring_swz(ra));
      fb[h][i] = rb * BK + 8 * (c ^ ring_swz(rb));
    }
  auto compute = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v8bf ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ah[i] = *(const v8bf*)&ring[buf][0][fa[h][i]];
        al[i] = *(const v8bf*)&ring[buf][1][fa[h][i]];
        bh[i] = *(const v8bf*)&ring[buf][2][fb[h][i]];
        bl[i] = *(const v8bf*)&ring[buf][3][fb[h][i]];
      }
      // the small products first, each accumulator four instructions apart
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
    }
  };
  const int nk = g.K / BK;
#pragma unroll
  for (int st = 0; st < NS - 1; ++st)
    if (st < nk) issue(st, st);
  for (int kt = 0; kt < nk; ++kt) {
    // stages kt+1 .. kt+ahead may stay in flight; stage kt must have landed
    const int ahead = min(NS - 2, nk - 1 - kt);
    if (ahead >= 2)
      wait_vmcnt<2 * PER_STAGE>();
    else if (ahead == 1)
      wait_vmcnt<PER_STAGE>();
    else
      wait_vmcnt<0>();
    asm volatile("s_barrier" ::: "memory");
    // the buffer of stage kt - 1 is free: every wave has passed its reads
    if (kt + NS - 1 < nk) issue(kt + NS - 1, (kt + NS - 1) % NS);
    compute(kt % NS);
  }
  // 32 x 32 accumulator, register r of lane l: m = l % 32, n = 8 (r / 4) + 4 (l / 32) + r % 4.
  // With beta the 16 old values of a tile are loaded together before any store.
  TO* C = (TO*)g.C;
  const bool with_beta = sizeof(TO) == 8 && g.beta != 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int m = m0 + wm + 32 * i + lr;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      TO* p[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) p[r] = C + m + (long)(n0 + wn + 32 * j + 8 * (r / 4) + 4 * lk + (r % 4)) * g.ldc;
      if (with_beta) {
        double old[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) old[r] = (double)*p[r];
#pragma unroll
        for (int r = 0; r < 16; ++r) *p[r] = (TO)((double)g.alpha * (double)acc[i][j][r] + g.beta * old[r]);
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) *p[r] = (TO)(g.alpha * acc[i][j][r]);
      }
    }
  }
}

// The MPO mix of H_AC on planes.  Tile: 32 consecutive a x 32 consecutive bb per block of 512
// threads (FP32 reads of whole 128-byte lines); thread t owns a = a0 + t % 32 and
// bb = bb0 + t / 32 + 16 r (r = 0, 1) with all w d input channels in registers:
//   in(i, s) = X0[a + D (s + d bb)] for level i = 0 (the identity level of GL), else
//              T1[a + D i + D w (s + d bb)] (FP32 output of GEMM 1)
//   out(u, j) = sum_{i,s} W[(u + d j) + w d (i + w s)] in(i, s), skipping all-zero d x d blocks
// Level j = w-1 (the identity level of GR) goes to Y[a + D (u + d bb)] (FP64); levels j < w-1
// are staged in LDS and written as planes Th/Tl[(a + D u) KT + D j + bb], KT = D (w - 1), with
// 4 lanes per 32-element row segment (the transpose that makes GEMM 2's A operand K-contiguous).
template <int MW, int MD>
__global__ void __launch_bounds__(512) wmix_planes_kernel(int D, int w, int d, const double* __restrict__ Wm,
                                                          const float* __restrict__ T1,
                                                          const double* __restrict__ X0, __bf16* __restrict__ Th,
                                                          __bf16* __restrict__ Tl, double* __restrict__ Y) {
  __shared__ float Ws[MW * MD * MW * MD];
  __shared__ int nzb[MW * MW];
  __shared__ float Lt[MD][32][33];
  const int wd = w * d;
  const int t = threadIdx.x;
  for (int e = t; e < wd * wd; e += 512) Ws[e] = (float)Wm[e];
  __syncthreads();
  if (t < w * w) {
    const int j = t / w, i = t % w;
    int nz = 0;
    for (int u = 0; u < d; ++u)
      for (int s = 0; s < d; ++s) nz |= Ws[(u + d * j) + wd * (i + w * s)] != 0.f;
    nzb[j * MW + i] = nz;
  }
  __syncthreads();
  const int a0 = blockIdx.x * 32, bb0 = blockIdx.y * 32;
  const int ta = t & 31, tb = t >> 5;
  const int a = a0 + ta;
  float in[2][MW][MD];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int bb = bb0 + tb + 16 * r;
#pragma unroll
    for (int i = 0; i < MW; ++i)
#pragma unroll
      for (int s = 0; s < MD; ++s) {
        in[r][i][s] = 0.f;
        if (i < w && s < d)
          in[r][i][s] = (i == 0) ? (float)X0[a + (long)D * (s + (long)d * bb)]
                                 : T1[a + (long)D * i + (long)D * w * (s + (long)d * bb)];
      }
  }
  const long KT = (long)D * (w - 1);
  // transposed write: row (a0 + ra, u = ru) of the tile, 8 bb from 8 q
  const int ra = (t >> 2) & 31, ru = t >> 7, q = t & 3;
  for (int j = 0; j < w; ++j) {
    float o[2][MD];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int u = 0; u < MD; ++u) o[r][u] = 0.f;
#pragma unroll
    for (int i = 0; i < MW; ++i) {
      if (i >= w || !nzb[j * MW + i]) continue;
#pragma unroll
      for (int s = 0; s < MD; ++s) {
        if (s >= d) continue;
#pragma unroll
        for (int u = 0; u < MD; ++u) {
          if (u >= d) continue;
          const float wv = Ws[(u + d * j) + wd * (i + w * s)];
          o[0][u] += wv * in[0][i][s];
          o[1][u] += wv * in[1][i][s];
        }
      }
    }
    if (j == w - 1) {
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int bb = bb0 + tb + 16 * r;
#pragma unroll
        for (int u = 0; u < MD; ++u)
          if (u < d) Y[a + (long)D * (u + (long)d * bb)] = (double)o[r][u];
      }
      continue;
    }
    __syncthreads();  // the previous level's reads of Lt are done
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int u = 0; u < MD; ++u)
        if (u < d) Lt[u][tb + 16 * r][ta] = o[r][u];
    __syncthreads();
    if (ru < d) {
      v8bf h, l;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 x, y;
        split_f32(Lt[ru][8 * q + e][ra], x, y);
        h[e] = x;
        l[e] = y;
      }
      const long off = (long)(a0 + ra + (long)D * ru) * KT + (long)D * j + bb0 + 8 * q;
      *(v8bf*)(Th + off) = h;
      *(v8bf*)(Tl + off) = l;
    }
  }
}

// y[a + D b] = sum_s P[s pstride + a + D b] + add[a + add_ld b]  (FP32 parts, FP64 sum)
__global__ void __launch_bounds__(256) planes_reduce_kernel(int D, int S, const float* __restrict__ P, long pstride,
                                                            const float* __restrict__ add, long add_ld,
                                                            double* __restrict__ y) {
  const int a = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
  if (a >= D) return;
  const long i = a + (long)D * b;
  double v = (double)add[a + add_ld * b];
  for (int z = 0; z < S; ++z) v += (double)P[z * pstride + i];
  y[i] = v;
}

bool al16(const void* p) { return ((uintptr_t)p & 15) == 0; }

}  // namespace

hipError_t launch_split_planes(int rows, int cols, const double* in, long ldi, bool trans, uint16_t* hi,
                               uint16_t* lo, long ldo, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  if (!al16(hi) || !al16(lo) || ldo % 8 != 0 || ldo < cols) return hipErrorInvalidValue;
  if (trans) {
    if (rows % 64 || cols % 64 || ldi < rows) return hipErrorInvalidValue;
    hipLaunchKernelGGL(split_trans_kernel<double>, dim3(rows / 64, cols / 64), dim3(256), 0, s, in, ldi, 0L,
                       (__bf16*)hi, (__bf16*)lo, ldo, 0L);
  } else {
    if (cols % 8 || ldi % 2 || ldi < cols || !al16(in)) return hipErrorInvalidValue;
    const long chunks = (long)rows * (cols / 8);
    const int blocks = (int)std::min<long>(4096, (chunks + 255) / 256);
    hipLaunchKernelGGL(split_rows_kernel, dim3(blocks), dim3(256), 0, s, rows, cols, in, ldi, (__bf16*)hi,
                       (__bf16*)lo, ldo);
  }
  return hipGetLastError();
}

hipError_t launch_gemm_planes(int M, int N, int K, const uint16_t* Ah, const uint16_t* Al, long lda,
                              const uint16_t* Bh, const uint16_t* Bl, long ldb, void* C, long ldc, bool c_double,
                              double alpha, double beta, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  PlaneGemm g{M, N, K, (const __bf16*)Ah, (const __bf16*)Al, lda, (const __bf16*)Bh, (const __bf16*)Bl, ldb,
              C, ldc, (float)alpha, beta, 0};
  constexpr int BK = 64;
  // full tiles, 16-byte chunks and 32-bit chunk offsets (rows of one tile times the stride)
  if (K % BK || lda % 8 || ldb % 8 || lda < K || ldb < K) return hipErrorInvalidValue;
  if (!al16(Ah) || !al16(Al) || !al16(Bh) || !al16(Bl)) return hipErrorInvalidValue;
  if (!c_double && beta != 0.0) return hipErrorInvalidValue;
  const long t128 = (long)(M / 128) * (N / 128);
  const bool full128 = M % 128 == 0 && N % 128 == 0;
  const bool big = full128 && t128 >= 512;
  // at most one workgroup per CU (long K on a small grid): the LDS-ring kernel keeps three
  // stages in flight
  const bool ring = full128 && t128 <= 256 && K >= 4 * 32;
  const int BM = 128, BN = (big || ring) ? 128 : 64;
  if (M % BM || N % BN || ldc < M) return hipErrorInvalidValue;
  if ((long)BM * lda * 2 >= (1L << 31) || (long)BN * ldb * 2 >= (1L << 31)) return hipErrorInvalidValue;
  dim3 grid(M / BM, N / BN);
  if (ring) {
    if (c_double)
      hipLaunchKernelGGL((gemm_planes_ring_kernel<4, double>), grid, dim3(256), 0, s, g);
    else
      hipLaunchKernelGGL((gemm_planes_ring_kernel<4, float>), grid, dim3(256), 0, s, g);
  } else if (big) {
    if (c_double)
      hipLaunchKernelGGL((gemm_planes_kernel<128, 128, BK, double>), grid, dim3(256), 0, s, g);
    else
      hipLaunchKernelGGL((gemm_planes_kernel<128, 128, BK, float>), grid, dim3(256), 0, s, g);
  } else {
    if (c_double)
      hipLaunchKernelGGL((gemm_planes_kernel<128, 64, BK, double>), grid, dim3(256), 0, s, g);
    else
      hipLaunchKernelGGL((gemm_planes_kernel<128, 64, BK, float>), grid, dim3(256), 0, s, g);
  }
  return hipGetLastError();
}

hipError_t launch_split_planes_levels(int rows, int cols, int nlev, const float* in, long ldi, long in_z,
                                      uint16_t* hi, uint16_t* lo, long ldo, long out_z, hipStream_t s) {
  if (rows <= 0 || cols <= 0 || nlev <= 0) return hipSuccess;
  if (rows % 64 || cols % 64 || ldi < rows || ldo < cols || ldo % 8 || out_z % 8) return hipErrorInvalidValue;
  if (!al16(hi) || !al16(lo)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(split_trans_kernel<float>, dim3(rows / 64, cols / 64, nlev), dim3(256), 0, s, in, ldi, in_z,
                     (__bf16*)hi, (__bf16*)lo, ldo, out_z);
  return hipGetLastError();
}

hipError_t launch_gemm_planes_splitk(int M, int N, int K, int S, const uint16_t* Ah, const uint16_t* Al, long lda,
                                     const uint16_t* Bh, const uint16_t* Bl, long ldb, float* P, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  constexpr int BK = 64;
  if (S < 1 || K % (S * BK) || M % 128 || N % 64) return hipErrorInvalidValue;
  if (lda % 8 || ldb % 8 || lda < K || ldb < K) return hipErrorInvalidValue;
  if (!al16(Ah) || !al16(Al) || !al16(Bh) || !al16(Bl)) return hipErrorInvalidValue;
  if (128L * lda * 2 >= (1L << 31) || 64L * ldb * 2 >= (1L << 31)) return hipErrorInvalidValue;
  PlaneGemm g{M, N, K / S, (const __bf16*)Ah, (const __bf16*)Al, lda, (const __bf16*)Bh, (const __bf16*)Bl, ldb,
              P, M, 1.0f, 0.0, (long)M * N};
  hipLaunchKernelGGL((gemm_planes_kernel<128, 64, BK, float>), dim3(M / 128, N / 64, S), dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t launch_planes_reduce(int D, int S, const float* P, long pstride, const float* add, long add_ld, double* y,
                                hipStream_t s) {
  if (D <= 0) return hipSuccess;
  hipLaunchKernelGGL(planes_reduce_kernel, dim3((D + 255) / 256, D), dim3(256), 0, s, D, S, P, pstride, add, add_ld,
                     y);
  return hipGetLastError();
}

hipError_t launch_wmix_planes(int D, int w, int d, const double* Wm, const float* T1, const double* X0,
                              uint16_t* Th, uint16_t* Tl, double* Y, hipStream_t s) {
  if (D % 32 || w < 2 || w > 8 || d > 4 || d < 1) return hipErrorInvalidValue;
  if (!al16(Th) || !al16(Tl)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((wmix_planes_kernel<8, 4>), dim3(D / 32, D / 32), dim3(512), 0, s, D, w, d, Wm, T1, X0,
                     (__bf16*)Th, (__bf16*)Tl, Y);
  return hipGetLastError();
}

}  // namespace imps
