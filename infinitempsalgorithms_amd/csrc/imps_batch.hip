// Device-batched VUMPS for parameter sweeps (one-site unit cells, uneven bond dimensions).
//
// Every instance runs the VUMPS loop of vumps.jl:51-86 (local update vumps.jl:7-19, gauge_step!
// vumps.jl:23-32 with uniform_rightorth! ortho.jl:166-202, recalculate! envs.jl:530-546 with the
// DMRGCache level order envs.jl:321-518, calc_galerkin vumps.jl:111-124, expectationvalue
// expval.jl:93-112), but the instances advance in lockstep: one host thread drives all of them
// and every kernel launch covers all active instances.  The Krylov solvers (Lanczos eigsolve,
// GMRES linsolve, Arnoldi gauge_fixedpoint) are per-instance state machines whose rounds are
// executed together: one grouped GEMM launch per contraction stage, one grouped Gram-Schmidt
// (CGS2) sequence and one host readback per Krylov step for all instances, whatever their bond
// dimensions.  QRpos/LQpos run as one batched Householder kernel (a workgroup per matrix, robust
// to the rank-deficient bond matrices of short-ranged states).
//
// The grouped kernels take arrays of per-instance descriptors (uploaded through a pinned arena
// with one copy per round); a workgroup finds its instance by binary search over the first
// global block of each descriptor.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "imps.h"
#include "imps_internal.h"

using namespace imps;

namespace {

struct BatchError : std::runtime_error {
  int code;
  BatchError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define BCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw BatchError(IMPS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
  } while (0)

void brequire(bool ok, int code, const std::string& msg) {
  if (!ok) throw BatchError(code, msg);
}

// ====================================================================== grouped kernels

// Operand pointers of the grouped kernels are read from descriptor arrays in device memory, so the
// compiler cannot tell that they address global memory; they are cast to the global address space
// (address space 1) to get global rather than flat loads and stores (flat accesses would also count
// against the LDS wait counter and serialise the LDS traffic of the tiled kernels).
typedef __attribute__((address_space(1))) double gdbl;
#define GP(p) ((gdbl*)(p))
#define GPC(p) ((const gdbl*)(p))

// Index of the item whose first block is the largest one <= b (items sorted by b0).
template <typename T>
__device__ inline int find_item(const T* __restrict__ items, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].b0 <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ---- grouped GEMM: C = alpha opA opB + beta C per item, 64 x 64 tiles, K stages of 16.
// opA(m, k) = A[m sAm + kk sAk + kc sAc], opB(k, n) = B[kk sBk + kc sBc + n sBn] with
// k = kc Kc + kk (Kc % 16 == 0 when Kc < K); either sAm or sAk is 1, either sBk or sBn is 1.
struct GG {
  int M, N, K, Kc;
  const double* A;
  long sAm, sAk, sAc;
  const double* B;
  long sBk, sBn, sBc;
  double* C;
  long ldc;
  double alpha, beta;
  int mtiles;
  int b0;  // first global tile of this item
};

template <typename TC>
struct BMfma;
template <>
struct BMfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct BMfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lk, int r) { return 4 * lk + r; }
};

template <typename TC>
__global__ void __launch_bounds__(256) ggemm_kernel(const GG* __restrict__ items, int nitems) {
  typedef BMfma<TC> MT;
  typedef typename MT::v4 v4;
  constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32, TM = 2, TN = 2, AE = 4, BE = 4;
  __shared__ TC As[BK][BM + 1];
  __shared__ TC Bs[BK][BN + 1];
  const GG g = items[find_item(items, nitems, (int)blockIdx.x)];
  const int loc = (int)blockIdx.x - g.b0;
  const int m0 = (loc % g.mtiles) * BM, n0 = (loc / g.mtiles) * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * WM, wn = (wave >> 1) * WN;
  const bool amc = g.sAm == 1, bkc = g.sBk == 1;
  v4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = v4{0, 0, 0, 0};
  const int lr = lane & 15, lk = lane >> 4;
  const int nkt = (g.K + BK - 1) / BK;
  for (int kt = 0; kt < nkt; ++kt) {
    const int kg0 = kt * BK;
    const int kc = kg0 / g.Kc, kk0 = kg0 - kc * g.Kc;
    const gdbl* Ab = GPC(g.A) + (long)kc * g.sAc + (long)kk0 * g.sAk;
    const gdbl* Bb = GPC(g.B) + (long)kc * g.sBc + (long)kk0 * g.sBk;
    double ra[AE], rb[BE];
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int idx = tid + 256 * e;
      const int m = amc ? idx % BM : idx / BK, k = amc ? idx / BM : idx % BK;
      ra[e] = (m0 + m < g.M && kg0 + k < g.K) ? Ab[(long)(m0 + m) * g.sAm + (long)k * g.sAk] : 0.0;
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int idx = tid + 256 * e;
      const int k = bkc ? idx % BK : idx / BN, n = bkc ? idx / BK : idx % BN;
      rb[e] = (n0 + n < g.N && kg0 + k < g.K) ? Bb[(long)k * g.sBk + (long)(n0 + n) * g.sBn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int idx = tid + 256 * e;
      const int m = amc ? idx % BM : idx / BK, k = amc ? idx / BM : idx % BK;
      As[k][m] = (TC)ra[e];
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int idx = tid + 256 * e;
      const int k = bkc ? idx % BK : idx / BN, n = bkc ? idx / BK : idx % BN;
      Bs[k][n] = (TC)rb[e];
    }
    __syncthreads();
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
  }
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    const int m = m0 + wm + i * 16 + lr;
    if (m >= g.M) continue;
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + j * 16 + MT::row(lk, r);
        if (n < g.N) {
          gdbl* p = GP(g.C) + m + (long)n * g.ldc;
          const double v = g.alpha * (double)acc[i][j][r];
          *p = (g.beta == 0.0) ? v : v + g.beta * *p;
        }
      }
  }
}

// ---- grouped vector operations (one opcode per item)
enum {
  GV_COPY = 0,   // y = x
  GV_AXPBY,      // y = a x + b y
  GV_ISQRT,      // y *= 1 / sqrt(*x) when *x > 0 (x: one device double)
  GV_FILL,       // y = a
  GV_MAXPY,      // y = b y + a sum_v coef[v] X[v ldX + i]  (b == 0: y not read)
  GV_COPY2D,     // y[r + ldy c] = x[r + ldx c], rows x cols
  GV_IDENT2D     // y[r + ldy c] = (r == c), rows x cols
};
struct GV {
  int op;
  int b0, nblk;
  long n;
  int rows;
  const double* x;
  double* y;
  long ldx, ldy;
  double a, b;
  const double* X;
  long ldX;
  int nv;
  const double* coef;
};

__global__ void __launch_bounds__(256) gvec_kernel(const GV* __restrict__ items, int nitems) {
  const GV it = items[find_item(items, nitems, (int)blockIdx.x)];
  const long stride = (long)it.nblk * 256;
  const long i0 = (long)((int)blockIdx.x - it.b0) * 256 + threadIdx.x;
  const gdbl* x = GPC(it.x);
  gdbl* y = GP(it.y);
  const gdbl* X = GPC(it.X);
  const gdbl* coef = GPC(it.coef);
  double sc = 0.0;
  if (it.op == GV_ISQRT) {
    const double s = *x;
    if (!(s > 0.0)) return;
    sc = 1.0 / sqrt(s);
  }
  for (long i = i0; i < it.n; i += stride) {
    switch (it.op) {
      case GV_COPY: y[i] = x[i]; break;
      case GV_AXPBY: y[i] = it.a * x[i] + it.b * y[i]; break;
      case GV_ISQRT: y[i] *= sc; break;
      case GV_FILL: y[i] = it.a; break;
      case GV_MAXPY: {
        double s = 0.0;
        for (int v = 0; v < it.nv; ++v) s += coef[v] * X[(long)v * it.ldX + i];
        y[i] = (it.b == 0.0 ? 0.0 : it.b * y[i]) + it.a * s;
        break;
      }
      case GV_COPY2D: {
        const long r = i % it.rows, c = i / it.rows;
        y[r + it.ldy * c] = x[r + it.ldx * c];
        break;
      }
      case GV_IDENT2D: {
        const long r = i % it.rows, c = i / it.rows;
        y[r + it.ldy * c] = (r == c) ? 1.0 : 0.0;
        break;
      }
      default: break;
    }
  }
}

// ---- grouped Gram-Schmidt (two classical passes, then the norm) and dot products.
// Item: w against V[0..nv) (ld n); partials part[v nb + blk]; coefficients coef[v] (first pass),
// coef[64 + v] (second pass), coef[128] (|w|^2 before normalisation).  A dot item has nv = 1,
// V = x, w = y and only runs the first pass (x . y into coef[0]).
struct GC {
  long n;
  int nv, nb, b0;
  const double* V;
  double* w;
  double* part;
  double* coef;
};

__device__ inline double bwave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int NV>
__device__ inline void bblock_store(double (&acc)[NV], int nvalid, gdbl* part, int stride, int blk) {
  __shared__ double red[NV][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    const double v = bwave_sum(acc[t]);
    if (lane == 0) red[t][wave] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < nvalid && threadIdx.x < NV)
    part[(size_t)threadIdx.x * stride + blk] = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                                               red[threadIdx.x][3];
  __syncthreads();
}

__global__ void __launch_bounds__(256) gc_dots_kernel(const GC* __restrict__ items, int nitems) {
  constexpr int CH = 32;
  const GC it = items[find_item(items, nitems, (int)blockIdx.x)];
  const int blk = (int)blockIdx.x - it.b0;
  for (int vc = 0; vc < it.nv; vc += CH) {
    const int nc = min(CH, it.nv - vc);
    const gdbl* Xc = GPC(it.V) + (size_t)vc * it.n;
    const gdbl* w = GPC(it.w);
    double acc[CH];
#pragma unroll
    for (int t = 0; t < CH; ++t) acc[t] = 0.0;
    for (long i = (long)blk * 256 + threadIdx.x; i < it.n; i += (long)it.nb * 256) {
      const double yi = w[i];
#pragma unroll
      for (int t = 0; t < CH; ++t)
        if (t < nc) acc[t] += Xc[(size_t)t * it.n + i] * yi;
    }
    bblock_store<CH>(acc, nc, GP(it.part) + (size_t)vc * it.nb, it.nb, blk);
  }
}

// w -= V c (c = coef + coff); DOTS: part <- V_v . w, else part <- w . w
template <bool DOTS>
__global__ void __launch_bounds__(256) gc_axpy_kernel(const GC* __restrict__ items, int nitems, int coff) {
  constexpr int MAXV = 32;
  constexpr int NA = DOTS ? MAXV : 1;
  const GC it = items[find_item(items, nitems, (int)blockIdx.x)];
  const int blk = (int)blockIdx.x - it.b0;
  __shared__ double c[MAXV];
  if (threadIdx.x < MAXV) c[threadIdx.x] = ((int)threadIdx.x < it.nv) ? GPC(it.coef)[coff + threadIdx.x] : 0.0;
  __syncthreads();
  const gdbl* V = GPC(it.V);
  gdbl* w = GP(it.w);
  double acc[NA];
#pragma unroll
  for (int t = 0; t < NA; ++t) acc[t] = 0.0;
  for (long i = (long)blk * 256 + threadIdx.x; i < it.n; i += (long)it.nb * 256) {
    double xv[MAXV];
    double wi = w[i];
#pragma unroll
    for (int v = 0; v < MAXV; ++v) {
      xv[v] = (v < it.nv) ? V[(size_t)v * it.n + i] : 0.0;
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
  bblock_store<NA>(acc, DOTS ? it.nv : 1, GP(it.part), it.nb, blk);
}

// coef[coff + v] = sum_b part[v nb + b] (nv values: nvals < 0 takes the item's nv); grid (items, 64)
__global__ void __launch_bounds__(64) gc_reduce_kernel(const GC* __restrict__ items, int coff, int nvals) {
  const GC it = items[blockIdx.x];
  const int v = blockIdx.y;
  const int nv = nvals < 0 ? it.nv : nvals;
  if (v >= nv) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < it.nb; b += 64) s += GPC(it.part)[(size_t)v * it.nb + b];
  s = bwave_sum(s);
  if (threadIdx.x == 0) GP(it.coef)[coff + v] = s;
}

__global__ void __launch_bounds__(256) gc_scale_kernel(const GC* __restrict__ items, int nitems) {
  const GC it = items[find_item(items, nitems, (int)blockIdx.x)];
  const double s = GPC(it.coef)[128];
  if (!(s > 0.0)) return;
  const double inv = 1.0 / sqrt(s);
  gdbl* w = GP(it.w);
  for (long i = (long)((int)blockIdx.x - it.b0) * 256 + threadIdx.x; i < it.n; i += (long)it.nb * 256) w[i] *= inv;
}

// ---- grouped H_AC W-mix (the launch_wmix contraction): T2[b,u,bb,j] = sum W[i,u,j,s] T1[b,i,s,bb],
// level i = 0 from X0 and channel j = w-1 to Y when given; blocks (ceil(D / 256) x D) per item
struct GW {
  int D, b0;
  const double* Wm;
  const double* T1;
  double* T2;
  const double* X0;
  double* Y;
};

template <int MAXIN>
__global__ void __launch_bounds__(256) gwmix_kernel(const GW* __restrict__ items, int nitems, int w, int d) {
  __shared__ double Ws[MAXIN * MAXIN];
  const GW it = items[find_item(items, nitems, (int)blockIdx.x)];
  const int loc = (int)blockIdx.x - it.b0;
  const int nbx = (it.D + 255) / 256;
  const int bb = loc / nbx;
  const int b = (loc % nbx) * 256 + threadIdx.x;
  const int wd = w * d, D = it.D;
  for (int t = threadIdx.x; t < wd * wd; t += blockDim.x) Ws[t] = GPC(it.Wm)[t];
  __syncthreads();
  if (b >= D) return;
  double x[MAXIN];
  const gdbl* src = GPC(it.T1) + b + (size_t)D * wd * bb;
  const gdbl* src0 = it.X0 ? GPC(it.X0) + b + (size_t)D * d * bb : nullptr;
#pragma unroll
  for (int q = 0; q < MAXIN; ++q)
    if (q < wd) {
      const int i = q % w;
      x[q] = (src0 && i == 0) ? src0[(size_t)D * (q / w)] : src[(size_t)D * q];
    }
  for (int j = 0; j < w; ++j) {
    gdbl* dst = (it.Y && j == w - 1) ? GP(it.Y) + b + (size_t)D * d * bb : GP(it.T2) + b + (size_t)D * d * (bb + (size_t)D * j);
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

// ---- grouped physical-index mix (slice pushes): out[a,u,b] = sum_s M[u,s] in[a,s,b] + beta out
struct GM {
  int D1, D2, b0;
  double M[64];
  const double* in;
  double* out;
  double beta;
};

__global__ void __launch_bounds__(256) gmix_kernel(const GM* __restrict__ items, int nitems, int d) {
  const int k = find_item(items, nitems, (int)blockIdx.x);
  const GM& it = items[k];
  const int loc = (int)blockIdx.x - it.b0;
  const int nbx = (it.D1 + 255) / 256;
  const int bcol = loc / nbx;
  const int a = (loc % nbx) * 256 + threadIdx.x;
  if (a >= it.D1) return;
  const size_t base = (size_t)a + (size_t)it.D1 * d * bcol;
  const gdbl* in = GPC(it.in);
  double x[8];
#pragma unroll
  for (int s = 0; s < 8; ++s)
    if (s < d) x[s] = in[base + (size_t)it.D1 * s];
  const double beta = it.beta;
  for (int u = 0; u < d; ++u) {
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < 8; ++s)
      if (s < d) acc += it.M[u + d * s] * x[s];
    gdbl* p = GP(it.out) + base + (size_t)it.D1 * u;
    *p = (beta == 0.0) ? acc : acc + beta * *p;
  }
}

// ---- grouped regularize! with the identity: v -= tr(l v) I (n x n; v leading dimension ldv,
// l contiguous and symmetric); one workgroup per item
struct GR {
  int n;
  long ldv;
  double* v;
  const double* l;
};

__global__ void __launch_bounds__(1024) greg_kernel(const GR* __restrict__ items) {
  const GR it = items[blockIdx.x];
  __shared__ double red[16];
  double s = 0.0;
  const long nn = (long)it.n * it.n;
  // tr(l v) = sum_{a,b} l[a + n b] v[b + ldv a]; l is symmetric here (C C^T or C^T C), so both
  // operands are read along their columns
  for (long e = threadIdx.x; e < nn; e += 1024) {
    const long b = e % it.n, a = e / it.n;
    s += GPC(it.l)[b + (long)it.n * a] * GPC(it.v)[b + it.ldv * a];
  }
  s = bwave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) t += red[k];
  for (int i = threadIdx.x; i < it.n; i += 1024) GP(it.v)[(long)i * it.ldv + i] -= t;
}

// ---- per-instance normalisation and change of a bond matrix (the end of a gauge sweep):
// C /= |C|, out = |C - Cpre|^2; one workgroup per item
struct GN {
  long n;
  double* C;
  const double* Cpre;
  double* out;
};

__global__ void __launch_bounds__(256) gnormdiff_kernel(const GN* __restrict__ items) {
  const GN it = items[blockIdx.x];
  __shared__ double red[4];
  __shared__ double bc;
  double s = 0.0;
  gdbl* C = GP(it.C);
  const gdbl* Cpre = GPC(it.Cpre);
  for (long i = threadIdx.x; i < it.n; i += 256) s += C[i] * C[i];
  s = bwave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = red[0] + red[1] + red[2] + red[3];
    bc = t > 0.0 ? 1.0 / sqrt(t) : 1.0;
  }
  __syncthreads();
  const double inv = bc;
  double e = 0.0;
  for (long i = threadIdx.x; i < it.n; i += 256) {
    const double c = C[i] * inv;
    C[i] = c;
    const double df = c - Cpre[i];
    e += df * df;
  }
  e = bwave_sum(e);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) *GP(it.out) = red[0] + red[1] + red[2] + red[3];
}

// ---- batched Householder QRpos / LQpos, one workgroup of 1024 threads per matrix.
// QR (trans = 0): X = in (mm x nn, leading dimension ldi, mm >= nn) = Q R; Q (mm x nn) to qout
// (ldq), R (nn x nn, upper, diagonal >= 0) to rout (ldr).
// LQ (trans = 1): the input is M (nn x mm, ldi) = L Q with X = M^T: Q (nn x mm) to qout, L (nn x
// nn, lower, diagonal >= 0) to rout.
// Reflectors as LAPACK's dlarfg (beta = -sign(x0) |x|, tau = (beta - x0) / beta, v0 = 1; tau = 0
// when the tail is zero), in panels of nb columns: a panel is factored column by column, its
// reflectors go to LDS, and every trailing column is then loaded once into the registers of one
// wavefront, receives the panel's nb reflectors there and is stored once (one pass over the
// trailing matrix per panel instead of per column).  Q is accumulated backwards panel by panel in
// the same way; finally each R row (L column) is made non-negative with its Q column (row).
// Scratch: W (mm x nn) and Q (mm x nn) per item.
struct GQ {
  int mm, nn, trans, want_q;
  const double* in;
  long ldi;
  double* qout;
  long ldq;
  double* rout;
  long ldr;
  double* W;
  double* Q;
};

constexpr int QR_THREADS = 1024;
constexpr int QR_MAXM = 2048;
constexpr int QR_LDS = 8192;  // doubles of panel reflectors (64 KiB)

__device__ inline double qr_block_sum(double v, double* red) {
  v = bwave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < QR_THREADS / 64; ++k) t += red[k];
  return t;
}

// Applies reflectors t = t0, t0 + dt, ... (count nt) of the panel starting at row k0 to the
// columns c0 and c1 (c1 may be null; rows k0..mm-1), held in registers of one wavefront (RPL rows
// per lane, RPL * 64 >= L).  Vs[t L + (i - k0)] = v_i of reflector t (zero above its own row),
// taus[t] its tau.  Two columns at a time keep two reduction chains in flight.
template <int RPL>
__device__ inline void qr_apply_panel(gdbl* c0, gdbl* c1, int k0, int L, const double* Vs, const double* taus,
                                      int t0, int nt, int dt) {
  const int lane = threadIdx.x & 63;
  double x[RPL], y[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int i = r * 64 + lane;
    x[r] = (i < L) ? c0[k0 + i] : 0.0;
    y[r] = (c1 && i < L) ? c1[k0 + i] : 0.0;
  }
  for (int c = 0, t = t0; c < nt; ++c, t += dt) {
    const double tk = taus[t];
    if (tk == 0.0) continue;
    const double* v = Vs + (size_t)t * L;
    double p = 0.0, q = 0.0;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int i = r * 64 + lane;
      const double vi = (i < L) ? v[i] : 0.0;
      p += vi * x[r];
      q += vi * y[r];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      p += __shfl_xor(p, off, 64);
      q += __shfl_xor(q, off, 64);
    }
    const double f = tk * p, g = tk * q;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int i = r * 64 + lane;
      const double vi = (i < L) ? v[i] : 0.0;
      x[r] -= f * vi;
      y[r] -= g * vi;
    }
  }
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    const int i = r * 64 + lane;
    if (i < L) {
      c0[k0 + i] = x[r];
      if (c1) c1[k0 + i] = y[r];
    }
  }
}

template <int RPL>
__global__ void __launch_bounds__(QR_THREADS) gqr_kernel(const GQ* __restrict__ items) {
  const GQ it = items[blockIdx.x];
  const int mm = it.mm, nn = it.nn;
  gdbl* W = GP(it.W);
  gdbl* Q = GP(it.Q);
  const gdbl* in = GPC(it.in);
  gdbl* qout = GP(it.qout);
  gdbl* rout = GP(it.rout);
  __shared__ double Vs[QR_LDS];
  __shared__ double tau[QR_MAXM / 2];
  __shared__ double red[QR_THREADS / 64];
  __shared__ double look[2];
  __shared__ double rdiag[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = QR_THREADS / 64;
  const int nb = max(1, min(16, QR_LDS / mm));
  const long tot = (long)mm * nn;
  for (long e = tid; e < tot; e += QR_THREADS) {
    const long i = e % mm, j = e / mm;
    W[e] = it.trans ? in[j + it.ldi * i] : in[i + it.ldi * j];
  }
  __syncthreads();
  for (int k0 = 0; k0 < nn; k0 += nb) {
    const int kb = min(nb, nn - k0);
    const int L = mm - k0;
    // the panel (rows k0..mm-1 of columns k0..k0+kb-1) is factored in LDS, column t at Vs + t L;
    // the wave that updates column t + 1 also leaves its new tail norm and diagonal entry in LDS
    for (int e = tid; e < kb * L; e += QR_THREADS) {
      const int t = e / L, i = e - t * L;
      Vs[e] = W[k0 + i + (long)mm * (k0 + t)];
    }
    __syncthreads();
    for (int t = 0; t < kb; ++t) {
      const int k = k0 + t;
      double* pc = Vs + (size_t)t * L;
      double tail, x0;
      if (t == 0) {
        double s = 0.0;
        for (int i = t + 1 + tid; i < L; i += QR_THREADS) s += pc[i] * pc[i];
        tail = qr_block_sum(s, red);
        x0 = pc[t];
      } else {
        tail = look[0];
        x0 = look[1];
      }
      double tk = 0.0, beta = x0, scal = 0.0;
      if (tail > 0.0) {
        const double nrm = sqrt(x0 * x0 + tail);
        beta = x0 >= 0.0 ? -nrm : nrm;
        tk = (beta - x0) / beta;
        scal = 1.0 / (x0 - beta);
      }
      __syncthreads();  // every thread has read x0 before row t becomes the reflector's 1
      for (int i = t + 1 + tid; i < L; i += QR_THREADS) pc[i] *= scal;
      if (tid == 0) {
        tau[k] = tk;
        rdiag[t] = beta;
        pc[t] = 1.0;
      }
      __syncthreads();
      for (int j = t + 1 + wave; j < kb; j += NW) {
        double* pj = Vs + (size_t)j * L;
        if (tk != 0.0) {
          double p = 0.0;
          for (int i = t + lane; i < L; i += 64) p += pc[i] * pj[i];
          p = __shfl(bwave_sum(p), 0, 64);
          const double f = tk * p;
          for (int i = t + lane; i < L; i += 64) pj[i] -= f * pc[i];
        }
        if (j == t + 1) {
          double q = 0.0;
          for (int i = t + 2 + lane; i < L; i += 64) q += pj[i] * pj[i];
          q = bwave_sum(q);
          if (lane == 0) {
            look[0] = q;
            look[1] = pj[t + 1];
          }
        }
      }
      __syncthreads();
    }
    // back to W (R above the diagonal, beta on it, reflectors below); then the R part in LDS is
    // cleared so that column t holds exactly reflector t for the trailing update
    for (int e = tid; e < kb * L; e += QR_THREADS) {
      const int t = e / L, i = e - t * L;
      W[k0 + i + (long)mm * (k0 + t)] = (i == t) ? rdiag[t] : Vs[e];
    }
    __syncthreads();
    for (int e = tid; e < kb * L; e += QR_THREADS) {
      const int t = e / L, i = e - t * L;
      if (i < t) Vs[e] = 0.0;
    }
    __syncthreads();
    // trailing columns: one pass each, the panel's reflectors applied in registers
    for (int j = k0 + kb + 2 * wave; j < nn; j += 2 * NW)
      qr_apply_panel<RPL>(W + (long)mm * j, j + 1 < nn ? W + (long)mm * (j + 1) : nullptr, k0, L, Vs, tau + k0, 0,
                          kb, 1);
    __syncthreads();
  }
  // Q = H_0 ... H_{nn-1} [I; 0], accumulated backwards panel by panel (skipped when only R is
  // wanted)
  const int np = it.want_q ? (nn + nb - 1) / nb : 0;
  for (long e = tid; it.want_q && e < tot; e += QR_THREADS) {
    const long i = e % mm, j = e / mm;
    Q[e] = (i == j) ? 1.0 : 0.0;
  }
  for (int pi = np - 1; pi >= 0; --pi) {
    const int k0 = pi * nb;
    const int kb = min(nb, nn - k0);
    const int L = mm - k0;
    __syncthreads();
    for (int t = 0; t < kb; ++t) {
      const int k = k0 + t;
      double* vt = Vs + (size_t)t * L;
      for (int i = tid; i < L; i += QR_THREADS) {
        const int row = k0 + i;
        vt[i] = (row < k) ? 0.0 : (row == k ? 1.0 : W[row + (long)mm * k]);
      }
    }
    __syncthreads();
    for (int j = k0 + 2 * wave; j < nn; j += 2 * NW)
      qr_apply_panel<RPL>(Q + (long)mm * j, j + 1 < nn ? Q + (long)mm * (j + 1) : nullptr, k0, L, Vs, tau + k0, kb - 1,
                          kb, -1);
  }
  __syncthreads();
  // signs of the diagonal of R, then the outputs
  for (int i = tid; i < nn; i += QR_THREADS) tau[i] = (W[i + (long)mm * i] < 0.0) ? -1.0 : 1.0;
  __syncthreads();
  for (long e = tid; it.want_q && e < tot; e += QR_THREADS) {
    const long i = e % mm, j = e / mm;
    const double q = Q[e] * tau[j];
    if (it.trans)
      qout[j + it.ldq * i] = q;
    else
      qout[i + it.ldq * j] = q;
  }
  const long rr = (long)nn * nn;
  for (long e = tid; e < rr; e += QR_THREADS) {
    const long i = e % nn, j = e / nn;
    const double r = (i <= j) ? W[i + (long)mm * j] * tau[i] : 0.0;
    if (it.trans)
      rout[j + it.ldr * i] = r;
    else
      rout[i + it.ldr * j] = r;
  }
}

// ---- batched Householder QRpos / LQpos with compact-WY trailing updates on FP64 MFMA
// (mm <= 512).  The matrix is kept row-major in the W scratch (ld nn), Q row-major in the Q
// scratch.  Per panel of 16 columns: the panel is factored in LDS exactly as in gqr_kernel, T of
// the compact form H_k0 ... H_k0+15 = I - V T V^T follows from G = V^T V (one MFMA product) and
// dlarft's recurrence, and every block of 16 trailing columns X gets X -= V (T^T (V^T X)) as two
// MFMA products (Q accumulation: Q -= V (T (V^T Q))).  Same reflectors and same results as
// gqr_kernel up to rounding.
constexpr int WY_NB = 16;
constexpr int WY_MAXM = 512;
constexpr int WY_LDV = WY_MAXM + 1;

typedef double wv4 __attribute__((ext_vector_type(4)));

__device__ inline wv4 wy_mma(double a, double b, wv4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// T (16 x 16, Ts[s * 16 + t]) of the panel's kb reflectors Vs (rows 0..L-1, reflector t at
// Vs[t * WY_LDV + r], columns kb..15 zero) and taus; called by every thread.  G = V^T V is summed
// from 16 per-wave MFMA partials (rows split over the waves, partials in Pw), then 16 lanes run
// dlarft's recurrence T[s][t] = -tau_t sum_{s <= q < t} T[s][q] G[q][t] with G's column t loaded
// ahead of each step.
__device__ inline void wy_make_t(const double* Vs, int L, int kb, const double* taus, double* Gs, double* Ts,
                                 double (*Pw)[256]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = QR_THREADS / 64;
  {
    wv4 acc = {0, 0, 0, 0};
    const int per = ((L + 4 * NW - 1) / (4 * NW)) * 4;  // rows per wave, a multiple of 4
    const int r0 = wave * per, r1 = min(L, r0 + per);
    for (int k4 = r0; k4 < r1; k4 += 4) {
      const int r = k4 + (lane >> 4);
      const double v = (r < r1) ? Vs[(lane & 15) * WY_LDV + r] : 0.0;
      acc = wy_mma(v, v, acc);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Pw[wave][((lane >> 4) + 4 * q) * 16 + (lane & 15)] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < 256) {
    double g = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) g += Pw[w][threadIdx.x];
    Gs[threadIdx.x] = g;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int s = threadIdx.x;
    double tr[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      double gc[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) gc[q] = Gs[q * 16 + t];
      double sum = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        if (q < t) sum += (q >= s ? tr[q] : 0.0) * gc[q];
      const double tt = (t < kb) ? taus[t] : 0.0;
      tr[t] = (s < t) ? -tt * sum : (s == t ? tt : 0.0);
    }
#pragma unroll
    for (int t = 0; t < 16; ++t) Ts[s * 16 + t] = tr[t];
  }
  __syncthreads();
}

// X[rows k0 + 0..L-1][columns c0..c0+15] -= V (T' (V^T X)) with T' = T^T (transT) or T, X row-major
// (ld nn, nn columns); one wavefront, Xw/Uw its 16 x 16 exchange buffers.  Rows go in chunks of 64:
// lane (lc, lg) holds rows base + 4 m + lg (m < 16) of column c0 + lc, which is both the MFMA B
// operand layout of V^T X and the accumulator layout of V U, so a chunk is loaded once per phase
// with its 16 loads in flight (the next chunk's loads are issued before this chunk's products).
__device__ inline void wy_load_chunk(const gdbl* X, int nn, int k0, int L, int base, int col, bool cok, int lg,
                                     double (&xb)[16]) {
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int r = base + 4 * m + lg;
    xb[m] = (cok && r < L) ? X[(long)(k0 + r) * nn + col] : 0.0;
  }
}

__device__ inline void wy_apply_block(gdbl* X, int nn, int mm, int k0, int L, int c0, const double* Vs,
                                      const double* Ts, bool transT, double* Xw, double* Uw) {
  const int lane = threadIdx.x & 63;
  const int lc = lane & 15, lg = lane >> 4;
  const int col = c0 + lc;
  const bool cok = col < nn;
  (void)mm;
  // Wt = V^T X (16 x 16)
  wv4 acc = {0, 0, 0, 0};
  double xb[16], xn[16];
#pragma unroll
  for (int m = 0; m < 16; ++m) xn[m] = 0.0;
  wy_load_chunk(X, nn, k0, L, 0, col, cok, lg, xb);
  for (int base = 0; base < L; base += 64) {
    if (base + 64 < L) wy_load_chunk(X, nn, k0, L, base + 64, col, cok, lg, xn);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int r = base + 4 * m + lg;
      const double v = (r < L) ? Vs[lc * WY_LDV + r] : 0.0;
      acc = wy_mma(v, xb[m], acc);
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) xb[m] = xn[m];
  }
  // lane holds Wt[t = lg + 4 q][c = lc]
#pragma unroll
  for (int q = 0; q < 4; ++q) Xw[(lg + 4 * q) * 16 + lc] = acc[q];
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the exchange is written
  __builtin_amdgcn_wave_barrier();
  // U = T' Wt: U[t][c] = sum_s T'[t][s] Wt[s][c]
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int t = lg + 4 * q;
    double u = 0.0;
#pragma unroll
    for (int s2 = 0; s2 < 16; ++s2) u += (transT ? Ts[s2 * 16 + t] : Ts[t * 16 + s2]) * Xw[s2 * 16 + lc];
    Uw[t * 16 + lc] = u;
  }
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  // B operand of V U for K step kk: U[4 kk + lg][lc]
  double bu[4];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) bu[kk] = Uw[(4 * kk + lg) * 16 + lc];
  // X -= V U, chunks of 64 rows (four row tiles of 16)
  wy_load_chunk(X, nn, k0, L, 0, col, cok, lg, xb);
  for (int base = 0; base < L; base += 64) {
    if (base + 64 < L) wy_load_chunk(X, nn, k0, L, base + 64, col, cok, lg, xn);
#pragma unroll
    for (int jt = 0; jt < 4; ++jt) {
      const int i0 = base + 16 * jt;
      wv4 d2 = {0, 0, 0, 0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int r = i0 + lc;
        const double a = (r < L) ? Vs[(4 * kk + lg) * WY_LDV + r] : 0.0;
        d2 = wy_mma(a, bu[kk], d2);
      }
      // d2 reg q: row i0 + lg + 4 q = base + 4 (4 jt + q) + lg, column lc -> xb[4 jt + q]
#pragma unroll
      for (int q = 0; q < 4; ++q) xb[4 * jt + q] -= d2[q];
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int r = base + 4 * m + lg;
      if (cok && r < L) X[(long)(k0 + r) * nn + col] = xb[m];
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) xb[m] = xn[m];
  }
}

__global__ void __launch_bounds__(QR_THREADS) gqr_wy_kernel(const GQ* __restrict__ items) {
  const GQ it = items[blockIdx.x];
  const int mm = it.mm, nn = it.nn;
  gdbl* W = GP(it.W);
  gdbl* Q = GP(it.Q);
  const gdbl* in = GPC(it.in);
  gdbl* qout = GP(it.qout);
  gdbl* rout = GP(it.rout);
  __shared__ double Vs[WY_NB * WY_LDV];
  __shared__ double Xs[QR_THREADS / 64][256];
  __shared__ double Us[QR_THREADS / 64][256];
  __shared__ double Ts[256];
  __shared__ double Gs[256];
  __shared__ double tau[WY_MAXM];
  __shared__ double red[QR_THREADS / 64];
  __shared__ double look[2];
  __shared__ double rdiag[WY_NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = QR_THREADS / 64;
  const long tot = (long)mm * nn;
  // W (row-major) = X: the LQ input M = X^T is read along its columns; the QR input is staged
  // through LDS tiles of 64 x 64 so that both sides stay coalesced
  if (it.trans) {
    for (long e = tid; e < tot; e += QR_THREADS) {
      const long i = e / nn, j = e - i * nn;
      W[e] = in[j + it.ldi * i];
    }
  } else {
    double* tile = Vs;  // 64 x 65
    for (int i0 = 0; i0 < mm; i0 += 64)
      for (int j0 = 0; j0 < nn; j0 += 64) {
        for (int e = tid; e < 4096; e += QR_THREADS) {
          const int i = e & 63, j = e >> 6;
          tile[i * 65 + j] = (i0 + i < mm && j0 + j < nn) ? in[(i0 + i) + it.ldi * (long)(j0 + j)] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < 4096; e += QR_THREADS) {
          const int j = e & 63, i = e >> 6;
          if (i0 + i < mm && j0 + j < nn) W[(long)(i0 + i) * nn + j0 + j] = tile[i * 65 + j];
        }
        __syncthreads();
      }
  }
  __syncthreads();
  for (int k0 = 0; k0 < nn; k0 += WY_NB) {
    const int kb = min(WY_NB, nn - k0);
    const int L = mm - k0;
    for (int e = tid; e < WY_NB * L; e += QR_THREADS) {
      const int r = e / WY_NB, t = e - r * WY_NB;
      Vs[t * WY_LDV + r] = (t < kb) ? W[(long)(k0 + r) * nn + k0 + t] : 0.0;
    }
    __syncthreads();
    for (int t = 0; t < kb; ++t) {
      const int k = k0 + t;
      double* pc = Vs + t * WY_LDV;
      double tail, x0;
      if (t == 0) {
        double s = 0.0;
        for (int i = t + 1 + tid; i < L; i += QR_THREADS) s += pc[i] * pc[i];
        tail = qr_block_sum(s, red);
        x0 = pc[t];
      } else {
        tail = look[0];
        x0 = look[1];
      }
      double tk = 0.0, beta = x0, scal = 0.0;
      if (tail > 0.0) {
        const double nrm = sqrt(x0 * x0 + tail);
        beta = x0 >= 0.0 ? -nrm : nrm;
        tk = (beta - x0) / beta;
        scal = 1.0 / (x0 - beta);
      }
      __syncthreads();
      for (int i = t + 1 + tid; i < L; i += QR_THREADS) pc[i] *= scal;
      if (tid == 0) {
        tau[k] = tk;
        rdiag[t] = beta;
        pc[t] = 1.0;
      }
      __syncthreads();
      for (int j = t + 1 + wave; j < kb; j += NW) {
        double* pj = Vs + j * WY_LDV;
        if (tk != 0.0) {
          double p = 0.0;
          for (int i = t + lane; i < L; i += 64) p += pc[i] * pj[i];
          p = __shfl(bwave_sum(p), 0, 64);
          const double f = tk * p;
          for (int i = t + lane; i < L; i += 64) pj[i] -= f * pc[i];
        }
        if (j == t + 1) {
          double q = 0.0;
          for (int i = t + 2 + lane; i < L; i += 64) q += pj[i] * pj[i];
          q = bwave_sum(q);
          if (lane == 0) {
            look[0] = q;
            look[1] = pj[t + 1];
          }
        }
      }
      __syncthreads();
    }
    // back to W; the R part of the panel in LDS is cleared (reflector t is zero above row t)
    for (int e = tid; e < kb * L; e += QR_THREADS) {
      const int r = e / kb, t = e - r * kb;
      W[(long)(k0 + r) * nn + k0 + t] = (r == t) ? rdiag[t] : Vs[t * WY_LDV + r];
    }
    __syncthreads();
    for (int e = tid; e < kb * L; e += QR_THREADS) {
      const int r = e / kb, t = e - r * kb;
      if (r < t) Vs[t * WY_LDV + r] = 0.0;
    }
    __syncthreads();
    if (k0 + kb < nn) {
      wy_make_t(Vs, L, kb, tau + k0, Gs, Ts, Xs);
      for (int c0 = k0 + kb + WY_NB * wave; c0 < nn; c0 += WY_NB * NW)
        wy_apply_block(W, nn, mm, k0, L, c0, Vs, Ts, true, Xs[wave], Us[wave]);
    }
    __syncthreads();
  }
  // Q = H_0 ... H_{nn-1} [I; 0], panels backwards: Q -= V (T (V^T Q)) on columns >= k0
  if (it.want_q) {
    for (long e = tid; e < tot; e += QR_THREADS) {
      const long i = e / nn, j = e - i * nn;
      Q[e] = (i == j) ? 1.0 : 0.0;
    }
    const int np = (nn + WY_NB - 1) / WY_NB;
    for (int pi = np - 1; pi >= 0; --pi) {
      const int k0 = pi * WY_NB;
      const int kb = min(WY_NB, nn - k0);
      const int L = mm - k0;
      __syncthreads();
      for (int e = tid; e < WY_NB * L; e += QR_THREADS) {
        const int r = e / WY_NB, t = e - r * WY_NB;
        Vs[t * WY_LDV + r] = (t >= kb || r < t) ? 0.0 : (r == t ? 1.0 : W[(long)(k0 + r) * nn + k0 + t]);
      }
      __syncthreads();
      wy_make_t(Vs, L, kb, tau + k0, Gs, Ts, Xs);
      for (int c0 = k0 + WY_NB * wave; c0 < nn; c0 += WY_NB * NW)
        wy_apply_block(Q, nn, mm, k0, L, c0, Vs, Ts, false, Xs[wave], Us[wave]);
    }
  }
  __syncthreads();
  for (int i = tid; i < nn; i += QR_THREADS) tau[i] = (W[(long)i * nn + i] < 0.0) ? -1.0 : 1.0;
  __syncthreads();
  if (it.want_q) {
    if (it.trans) {
      for (long e = tid; e < tot; e += QR_THREADS) {
        const long i = e / nn, j = e - i * nn;
        qout[j + it.ldq * i] = Q[e] * tau[j];
      }
    } else {
      double* tile = Vs;
      for (int i0 = 0; i0 < mm; i0 += 64)
        for (int j0 = 0; j0 < nn; j0 += 64) {
          for (int e = tid; e < 4096; e += QR_THREADS) {
            const int j = e & 63, i = e >> 6;
            tile[i * 65 + j] = (i0 + i < mm && j0 + j < nn) ? Q[(long)(i0 + i) * nn + j0 + j] * tau[j0 + j] : 0.0;
          }
          __syncthreads();
          for (int e = tid; e < 4096; e += QR_THREADS) {
            const int i = e & 63, j = e >> 6;
            if (i0 + i < mm && j0 + j < nn) qout[(i0 + i) + it.ldq * (long)(j0 + j)] = tile[i * 65 + j];
          }
          __syncthreads();
        }
    }
  }
  const long rr = (long)nn * nn;
  for (long e = tid; e < rr; e += QR_THREADS) {
    const long i = e / nn, j = e - i * nn;
    const double r = (i <= j) ? W[i * nn + j] * tau[i] : 0.0;
    if (it.trans)
      rout[j + it.ldr * i] = r;
    else
      rout[i + it.ldr * j] = r;
  }
}

}  // namespace

// ====================================================================== host side

namespace {

constexpr int kCoef = 129;  // per Gram-Schmidt item: two passes of 64 coefficients and |w|^2
constexpr int kMaxBasis = 32;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Pinned host staging mirrored by a device buffer: descriptors and coefficients of one round are
// written on the host, copied with one transfer, and read by the kernels from the device copy.
// Reset only after the stream has been synchronised.
struct Arena {
  char* h = nullptr;
  char* d = nullptr;
  size_t cap = 0, off = 0, flushed = 0;
  void init(size_t bytes) {
    cap = bytes;
    BCHK(hipHostMalloc((void**)&h, cap, hipHostMallocDefault));
    BCHK(hipMalloc((void**)&d, cap));
  }
  void release() {
    if (h) (void)hipHostFree(h);
    if (d) (void)hipFree(d);
    h = d = nullptr;
  }
  // copies n objects into the staging area; returns their future device address
  template <typename T>
  T* put(const T* src, size_t n) {
    const size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
    off = align_up(off, 256);
    brequire(off + bytes <= cap, IMPS_E_ARG, "batch: descriptor arena exhausted");
    if (n) std::memcpy(h + off, src, sizeof(T) * n);
    T* dev = (T*)(d + off);
    off += bytes;
    return dev;
  }
  void flush(hipStream_t s) {
    if (off > flushed) BCHK(hipMemcpyAsync(d + flushed, h + flushed, off - flushed, hipMemcpyHostToDevice, s));
    flushed = off;
  }
  void reset() { off = flushed = 0; }
};

// Device scratch for reduction partials and results of one round (bump allocation, reset per round)
struct Scratch {
  double* d = nullptr;
  size_t cap = 0, off = 0;
  double* take(size_t n) {
    off = align_up(off, 32);
    brequire(off + n <= cap, IMPS_E_ARG, "batch: scratch exhausted");
    double* p = d + off;
    off += n;
    return p;
  }
};

enum OpKind { OP_HAC, OP_HACFULL, OP_HC, OP_PUSHL, OP_PUSHR, OP_SLICEL, OP_SLICER };

struct Inst;

// A linear map of one instance applied to device vectors: H_AC (identity structural levels applied
// as copies, effective.jl:90-102), the first stage of H_AC with every level (energy), H_C
// (effective.jl:72-76), identity-channel transfer pushes with distinct layers (transfermatrix.jl:
// 122-191) optionally followed by regularize! with the identity (envs.jl:28-32), and slice pushes
// with one d x d block (envs.jl:104-114).
struct Op {
  OpKind kind = OP_HAC;
  Inst* in = nullptr;
  int prec = PREC_FP64;
  const double* above = nullptr;
  const double* below = nullptr;
  const double* rho = nullptr;
  double M[64] = {};
};

struct Inst {
  int idx = 0, D = 0, d = 0, w = 0;
  size_t nA = 0, nC = 0, nE = 0;
  std::vector<double> Wh;  // (w, d, w, d)
  double* mem = nullptr;
  double *Wm, *AL, *AR, *C, *AC, *pAC, *pC, *pAL, *GL, *GR, *oldGL, *oldGR;
  double *T1, *T2, *T1c, *Y, *Z, *Ypc, *Zpc;
  double *VA, *tmpA, *VC, *tmpC, *VG, *gt, *gr, *gx, *grhs, *VR, *ev;
  double *qW, *qQ, *qW2, *qQ2, *Qa, *Qc, *Rq, *Rq2, *Cpre, *Mg, *Qdump, *gk, *gq, *Id, *rho;
  int kdA = 30, kdC = 30, kdE = 30;
  // loop state
  bool active = true, converged = false;
  int status = IMPS_OK;
  int niter = 0, prec = PREC_FP64, switch_iter = -1, sweeps = 0;
  double eps = 0.0, best = std::numeric_limits<double>::infinity(), f = 0.0, lam_ac = 1.0, lam_c = 1.0;
  bool stalled = false;
  long hac_applies = 0, gmres_ops = 0;
  double seconds = 0.0;
  // environments' warm start available
  bool envs_valid = false;
};

// one step of a phase: a GEMM or a vector operation, with optional uploaded coefficients
struct Item {
  int kind = 0;  // 0 GEMM, 1 vector op
  int prec = PREC_FP64;
  GG g{};
  GV v{};
  std::vector<double> upload;
  int bind = 0;  // 1: v.coef, 2: g.B point at the uploaded coefficients
};

Item gemm_item(int M, int N, int K, double alpha, const double* A, long sAm, long sAk, const double* B, long sBk,
               long sBn, double beta, double* C, long ldc, int prec = PREC_FP64) {
  Item it;
  it.kind = 0;
  it.prec = prec;
  it.g.M = M;
  it.g.N = N;
  it.g.K = K;
  it.g.Kc = K;
  it.g.A = A;
  it.g.sAm = sAm;
  it.g.sAk = sAk;
  it.g.B = B;
  it.g.sBk = sBk;
  it.g.sBn = sBn;
  it.g.C = C;
  it.g.ldc = ldc;
  it.g.alpha = alpha;
  it.g.beta = beta;
  return it;
}

Item vec_item(int op, long n, const double* x, double* y, double a = 0.0, double b = 0.0) {
  Item it;
  it.kind = 1;
  it.v.op = op;
  it.v.n = n;
  it.v.x = x;
  it.v.y = y;
  it.v.a = a;
  it.v.b = b;
  return it;
}

Item copy2d_item(int rows, int cols, const double* x, long ldx, double* y, long ldy) {
  Item it = vec_item(GV_COPY2D, (long)rows * cols, x, y);
  it.v.rows = rows;
  it.v.ldx = ldx;
  it.v.ldy = ldy;
  return it;
}

Item ident_item(int n, double* y, long ldy) {
  Item it = vec_item(GV_IDENT2D, (long)n * n, nullptr, y);
  it.v.rows = n;
  it.v.ldy = ldy;
  return it;
}

// y = b y + a sum_v c[v] X[v ldX]
Item maxpy_item(long n, const double* X, long ldX, const std::vector<double>& c, double a, double b, double* y) {
  Item it = vec_item(GV_MAXPY, n, nullptr, y, a, b);
  it.v.X = X;
  it.v.ldX = ldX;
  it.v.nv = (int)c.size();
  it.upload = c;
  it.bind = 1;
  return it;
}

// The device work of one problem for one round (all parts optional, executed in this order):
// pre items, one operator application src -> dst, post items, Gram-Schmidt of w against V[0..nv)
// with normalisation, dot products.  After the round coef points at the Gram-Schmidt results
// (host copy) and dotv holds the dots.
struct Phase {
  std::vector<Item> pre, post;
  const Op* op = nullptr;
  const double* src = nullptr;
  double* dst = nullptr;
  bool cgs = false;
  const double* V = nullptr;
  int nv = 0;
  double* w = nullptr;
  long n = 0;
  std::vector<const double*> dx, dy;
  std::vector<long> dn;
  // results
  const double* coef = nullptr;
  std::vector<double> dotv;
  // device result slots (set while executing)
  double* dcoef = nullptr;
  double* ddots = nullptr;
  void dot(const double* x, const double* y, long nn) {
    dx.push_back(x);
    dy.push_back(y);
    dn.push_back(nn);
  }
};

int red_blocks(long n) { return (int)std::min<long>(64, std::max<long>(1, (n + 2047) / 2048)); }
int vec_blocks(long n) { return (int)std::min<long>(512, std::max<long>(1, (n + 1023) / 1024)); }

// Host threads for the per-instance bookkeeping of a round (projected eigenproblems, Givens
// rotations): run(n, f) calls f(0..n-1) spread over the pool and the calling thread.
struct Pool {
  std::vector<std::thread> th;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  const std::function<void(int)>* job = nullptr;
  int n = 0, busy = 0;
  long gen = 0;
  bool stop = false;
  std::atomic<int> next{0};
  void start(int nt) {
    for (int i = 0; i < nt; ++i) th.emplace_back([this] { loop(); });
  }
  void work() {
    int k;
    while ((k = next.fetch_add(1)) < n) (*job)(k);
  }
  void loop() {
    long seen = 0;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || gen != seen; });
        if (stop) return;
        seen = gen;
      }
      work();
      std::lock_guard<std::mutex> lk(mu);
      if (--busy == 0) cv_done.notify_all();
    }
  }
  void run(int count, const std::function<void(int)>& f) {
    if (count < 4 || th.empty()) {
      for (int k = 0; k < count; ++k) f(k);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      job = &f;
      n = count;
      next = 0;
      busy = (int)th.size();
      gen++;
    }
    cv.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu);
    cv_done.wait(lk, [&] { return busy == 0; });
  }
  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    for (auto& t : th) t.join();
    th.clear();
  }
};

struct Batch {
  int device = 0;
  hipStream_t s = nullptr;
  Pool pool;
  // kernel launches of the current round, issued by submit() after one upload of the round's
  // descriptors (no host reads happen between them)
  std::vector<std::function<void()>> pending;
  Arena ar;
  Scratch sc;
  double* hres = nullptr;  // pinned readback
  size_t hres_cap = 0;
  int w = 0, d = 0;
  long launches = 0, rounds = 0;
  // wall seconds per phase: canonicalisation, local update (eigensolves + regauge!), gauge_step!,
  // recalculate!, calc_galerkin, expectationvalue
  double t_phase[6] = {0, 0, 0, 0, 0, 0};
  std::vector<Inst*> insts;

  void init(int dev, int pool_threads) {
    device = dev;
    BCHK(hipSetDevice(dev));
    BCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    ar.init((size_t)32 << 20);
    sc.cap = (size_t)4 << 20;
    BCHK(hipMalloc((void**)&sc.d, sc.cap * sizeof(double)));
    hres_cap = (size_t)1 << 20;
    BCHK(hipHostMalloc((void**)&hres, hres_cap * sizeof(double), hipHostMallocDefault));
    pool.start(pool_threads);
  }
  void release() {
    pool.shutdown();
    if (s) (void)hipStreamSynchronize(s);
    for (Inst* in : insts) {
      if (in->mem) (void)hipFree(in->mem);
      delete in;
    }
    insts.clear();
    ar.release();
    if (sc.d) (void)hipFree(sc.d);
    if (hres) (void)hipHostFree(hres);
    if (s) (void)hipStreamDestroy(s);
    s = nullptr;
  }
  void enqueue(std::function<void()> f) { pending.push_back(std::move(f)); }
  // one descriptor upload, then the round's launches in order
  void submit() {
    ar.flush(s);
    for (auto& f : pending) {
      f();
      BCHK(hipGetLastError());
    }
    launches += (long)pending.size();
    pending.clear();
  }
  void sync() {
    submit();
    BCHK(hipStreamSynchronize(s));
    ar.reset();
    sc.off = 0;
  }

  // ---- grouped launches (descriptors already in the arena unless noted)
  void launch_gemms(std::vector<GG>& gs, int prec) {
    if (gs.empty()) return;
    int t = 0;
    for (GG& g : gs) {
      g.mtiles = (g.M + 63) / 64;
      g.b0 = t;
      t += g.mtiles * ((g.N + 63) / 64);
    }
    const GG* dg = ar.put(gs.data(), gs.size());
    if (t == 0) return;
    const int ng = (int)gs.size();
    hipStream_t st = s;
    if (prec == PREC_FP32)
      enqueue([=] { hipLaunchKernelGGL(ggemm_kernel<float>, dim3(t), dim3(256), 0, st, dg, ng); });
    else
      enqueue([=] { hipLaunchKernelGGL(ggemm_kernel<double>, dim3(t), dim3(256), 0, st, dg, ng); });
  }
  void launch_vecs(std::vector<GV>& vs) {
    if (vs.empty()) return;
    int t = 0;
    for (GV& v : vs) {
      v.nblk = vec_blocks(v.n);
      v.b0 = t;
      t += v.nblk;
    }
    const GV* dv = ar.put(vs.data(), vs.size());
    const int nv = (int)vs.size();
    hipStream_t st = s;
    enqueue([=] { hipLaunchKernelGGL(gvec_kernel, dim3(t), dim3(256), 0, st, dv, nv); });
  }
  // items at one position of several problems: GEMMs grouped per precision, vector ops together
  void run_items(std::vector<Item*>& items) {
    std::vector<GG> g64, g32;
    std::vector<GV> vs;
    for (Item* it : items) {
      if (!it->upload.empty()) {
        double* dp = ar.put(it->upload.data(), it->upload.size());
        if (it->bind == 1) it->v.coef = dp;
        if (it->bind == 2) it->g.B = dp;
      }
      if (it->kind == 0) {
        if (it->g.M > 0 && it->g.N > 0) (it->prec == PREC_FP32 ? g32 : g64).push_back(it->g);
      } else if (it->v.n > 0) {
        vs.push_back(it->v);
      }
    }
    launch_gemms(g64, PREC_FP64);
    launch_gemms(g32, PREC_FP32);
    launch_vecs(vs);
  }

  void apply_ops(const std::vector<Phase*>& ph);
  void execute(const std::vector<Phase*>& ph);
};

// Operator applications of all phases: stage 1 GEMMs, then the mixes and copies, stage 2 GEMMs,
// then regularize!.  Every instance has at most one application per operator kind in a round
// (the temporaries are per instance and kind).
void Batch::apply_ops(const std::vector<Phase*>& ph) {
  std::vector<Item> s1, s3;
  std::vector<GW> wm;
  std::vector<GM> mx;
  std::vector<GV> cp;
  std::vector<GR> rg;
  for (Phase* p : ph) {
    if (!p->op) continue;
    const Op& o = *p->op;
    Inst& I = *o.in;
    const int D = I.D, d = I.d, w = I.w;
    const double* x = p->src;
    double* y = p->dst;
    switch (o.kind) {
      case OP_HAC: {
        // T1 = GL[levels 1..w-1] x; mix (level 0 from x, channel w-1 into y); y += T2 GR
        s1.push_back(gemm_item(D * (w - 1), d * D, D, 1.0, I.GL + D, 1, (long)D * w, x, 1, D, 0.0, I.T1 + D,
                               (long)D * w, o.prec));
        GW g{};
        g.D = D;
        g.Wm = I.Wm;
        g.T1 = I.T1;
        g.T2 = I.T2;
        g.X0 = x;
        g.Y = y;
        wm.push_back(g);
        s3.push_back(gemm_item(D * d, D, D * (w - 1), 1.0, I.T2, 1, (long)D * d, I.GR, 1, (long)D * w, 1.0, y,
                               (long)D * d, o.prec));
        I.hac_applies++;
        break;
      }
      case OP_HACFULL: {
        // T1 = GL x with every level, T2 = W-mix (channel w-1 is <AC| H_AC's last row>)
        s1.push_back(gemm_item(D * w, d * D, D, 1.0, I.GL, 1, (long)D * w, x, 1, D, 0.0, I.T1, (long)D * w));
        GW g{};
        g.D = D;
        g.Wm = I.Wm;
        g.T1 = I.T1;
        g.T2 = I.T2;
        wm.push_back(g);
        break;
      }
      case OP_HC: {
        // T1c[levels 1..w-1] = GL x, T1c[0] = x, y = T1c[w-1] + sum_{k < w-1} T1c[k] GR[k]
        s1.push_back(gemm_item(D * (w - 1), D, D, 1.0, I.GL + D, 1, (long)D * w, x, 1, D, 0.0, I.T1c + D, (long)D * w,
                               o.prec));
        GV c0{};
        c0.op = GV_COPY2D;
        c0.n = (long)D * D;
        c0.rows = D;
        c0.x = x;
        c0.ldx = D;
        c0.y = I.T1c;
        c0.ldy = (long)D * w;
        cp.push_back(c0);
        GV c1 = c0;
        c1.x = I.T1c + (size_t)D * (w - 1);
        c1.ldx = (long)D * w;
        c1.y = y;
        c1.ldy = D;
        cp.push_back(c1);
        Item g2 = gemm_item(D, D, D * (w - 1), 1.0, I.T1c, 1, (long)D * w, I.GR, 1, (long)D * w, 1.0, y, D, o.prec);
        g2.g.Kc = D;
        g2.g.sAc = D;
        g2.g.sBc = D;
        s3.push_back(g2);
        break;
      }
      case OP_PUSHL:
      case OP_SLICEL: {
        // out[a',b'] = sum below[a,s,a'] (M) X[a,b] above[b,s,b']
        s1.push_back(gemm_item(D, d * D, D, 1.0, x, 1, D, o.above, 1, D, 0.0, I.Ypc, D));
        const double* src = I.Ypc;
        if (o.kind == OP_SLICEL) {
          GM m{};
          m.D1 = D;
          m.D2 = D;
          std::memcpy(m.M, o.M, sizeof(m.M));
          m.in = I.Ypc;
          m.out = I.Zpc;
          mx.push_back(m);
          src = I.Zpc;
        }
        s3.push_back(gemm_item(D, D, d * D, 1.0, o.below, (long)d * D, 1, src, 1, (long)d * D, 0.0, y, D));
        break;
      }
      case OP_PUSHR:
      case OP_SLICER: {
        // out[a',b'] = sum above[a',s,a] X[a,b] (M) below[b',s,b]
        s1.push_back(gemm_item(d * D, D, D, 1.0, o.above, 1, (long)d * D, x, 1, D, 0.0, I.Ypc, (long)d * D));
        const double* src = I.Ypc;
        if (o.kind == OP_SLICER) {
          GM m{};
          m.D1 = D;
          m.D2 = D;
          std::memcpy(m.M, o.M, sizeof(m.M));
          m.in = I.Ypc;
          m.out = I.Zpc;
          mx.push_back(m);
          src = I.Zpc;
        }
        s3.push_back(gemm_item(D, D, d * D, 1.0, src, 1, D, o.below, D, 1, 0.0, y, D));
        break;
      }
    }
    if (o.rho && (o.kind == OP_PUSHL || o.kind == OP_PUSHR)) {
      GR r{};
      r.n = D;
      r.ldv = D;
      r.v = y;
      r.l = o.rho;
      rg.push_back(r);
    }
  }
  {
    std::vector<Item*> it;
    for (Item& x : s1) it.push_back(&x);
    run_items(it);
  }
  if (!wm.empty()) {
    int t = 0;
    for (GW& g : wm) {
      g.b0 = t;
      t += ((g.D + 255) / 256) * g.D;
    }
    const GW* dg = ar.put(wm.data(), wm.size());
    const int wd = w * d, ng = (int)wm.size(), ww = w, dd = d;
    hipStream_t st = s;
    if (wd <= 16)
      enqueue([=] { hipLaunchKernelGGL(gwmix_kernel<16>, dim3(t), dim3(256), 0, st, dg, ng, ww, dd); });
    else if (wd <= 32)
      enqueue([=] { hipLaunchKernelGGL(gwmix_kernel<32>, dim3(t), dim3(256), 0, st, dg, ng, ww, dd); });
    else
      enqueue([=] { hipLaunchKernelGGL(gwmix_kernel<64>, dim3(t), dim3(256), 0, st, dg, ng, ww, dd); });
  }
  if (!mx.empty()) {
    int t = 0;
    for (GM& m : mx) {
      m.b0 = t;
      t += ((m.D1 + 255) / 256) * m.D2;
    }
    const GM* dm = ar.put(mx.data(), mx.size());
    const int ng = (int)mx.size(), dd = d;
    hipStream_t st = s;
    enqueue([=] { hipLaunchKernelGGL(gmix_kernel, dim3(t), dim3(256), 0, st, dm, ng, dd); });
  }
  launch_vecs(cp);
  {
    std::vector<Item*> it;
    for (Item& x : s3) it.push_back(&x);
    run_items(it);
  }
  if (!rg.empty()) {
    const GR* dr = ar.put(rg.data(), rg.size());
    const unsigned ng = (unsigned)rg.size();
    hipStream_t st = s;
    enqueue([=] { hipLaunchKernelGGL(greg_kernel, dim3(ng), dim3(1024), 0, st, dr); });
  }
}

void Batch::execute(const std::vector<Phase*>& ph) {
  rounds++;
  size_t npre = 0, npost = 0;
  for (Phase* p : ph) {
    npre = std::max(npre, p->pre.size());
    npost = std::max(npost, p->post.size());
  }
  for (size_t k = 0; k < npre; ++k) {
    std::vector<Item*> it;
    for (Phase* p : ph)
      if (k < p->pre.size()) it.push_back(&p->pre[k]);
    run_items(it);
  }
  apply_ops(ph);
  for (size_t k = 0; k < npost; ++k) {
    std::vector<Item*> it;
    for (Phase* p : ph)
      if (k < p->post.size()) it.push_back(&p->post[k]);
    run_items(it);
  }
  // result slots: Gram-Schmidt coefficients, then dots, contiguous for one readback
  size_t nres = 0;
  for (Phase* p : ph) nres += (p->cgs ? kCoef : 0) + p->dx.size();
  double* res = nres ? sc.take(nres) : nullptr;
  size_t off = 0;
  std::vector<GC> gcs, gds;
  for (Phase* p : ph) {
    if (p->cgs) {
      brequire(p->nv >= 1 && p->nv <= kMaxBasis, IMPS_E_ARG, "batch: Gram-Schmidt basis size out of range");
      GC g{};
      g.n = p->n;
      g.nv = p->nv;
      g.nb = red_blocks(p->n);
      g.V = p->V;
      g.w = p->w;
      g.part = sc.take((size_t)p->nv * g.nb);
      g.coef = res + off;
      p->dcoef = res + off;
      off += kCoef;
      gcs.push_back(g);
    }
  }
  for (Phase* p : ph) {
    p->ddots = res + off;
    for (size_t k = 0; k < p->dx.size(); ++k) {
      GC g{};
      g.n = p->dn[k];
      g.nv = 1;
      g.nb = red_blocks(p->dn[k]);
      g.V = p->dx[k];
      g.w = const_cast<double*>(p->dy[k]);
      g.part = sc.take(g.nb);
      g.coef = res + off;
      off += 1;
      gds.push_back(g);
    }
  }
  auto prefix = [](std::vector<GC>& v) {
    int t = 0;
    for (GC& g : v) {
      g.b0 = t;
      t += g.nb;
    }
    return t;
  };
  if (!gcs.empty()) {
    const int t = prefix(gcs);
    int maxnv = 1;
    for (const GC& g : gcs) maxnv = std::max(maxnv, g.nv);
    const GC* dg = ar.put(gcs.data(), gcs.size());
    const int ni = (int)gcs.size();
    hipStream_t st = s;
    enqueue([=] { hipLaunchKernelGGL(gc_dots_kernel, dim3(t), dim3(256), 0, st, dg, ni); });
    enqueue([=] { hipLaunchKernelGGL(gc_reduce_kernel, dim3(ni, maxnv), dim3(64), 0, st, dg, 0, -1); });
    enqueue([=] { hipLaunchKernelGGL(gc_axpy_kernel<true>, dim3(t), dim3(256), 0, st, dg, ni, 0); });
    enqueue([=] { hipLaunchKernelGGL(gc_reduce_kernel, dim3(ni, maxnv), dim3(64), 0, st, dg, 64, -1); });
    enqueue([=] { hipLaunchKernelGGL(gc_axpy_kernel<false>, dim3(t), dim3(256), 0, st, dg, ni, 64); });
    enqueue([=] { hipLaunchKernelGGL(gc_reduce_kernel, dim3(ni, 1), dim3(64), 0, st, dg, 128, 1); });
    enqueue([=] { hipLaunchKernelGGL(gc_scale_kernel, dim3(t), dim3(256), 0, st, dg, ni); });
  }
  if (!gds.empty()) {
    const int t = prefix(gds);
    const GC* dg = ar.put(gds.data(), gds.size());
    const int ni = (int)gds.size();
    hipStream_t st = s;
    enqueue([=] { hipLaunchKernelGGL(gc_dots_kernel, dim3(t), dim3(256), 0, st, dg, ni); });
    enqueue([=] { hipLaunchKernelGGL(gc_reduce_kernel, dim3(ni, 1), dim3(64), 0, st, dg, 0, 1); });
  }
  submit();
  if (nres) {
    brequire(nres <= hres_cap, IMPS_E_ARG, "batch: readback too large");
    BCHK(hipMemcpyAsync(hres, res, nres * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  sync();
  off = 0;
  for (Phase* p : ph) {
    if (p->cgs) {
      p->coef = hres + off;
      off += kCoef;
    }
  }
  for (Phase* p : ph) {
    p->dotv.assign(hres + off, hres + off + p->dx.size());
    off += p->dx.size();
  }
}

}  // namespace

namespace {

// A Krylov solve of one instance, advanced one round at a time: next() describes the device work
// of the round, consume() reads its results.
struct Prob {
  bool done = false;
  int status = IMPS_OK;
  virtual ~Prob() {}
  virtual void next(Phase& ph) = 0;
  virtual void consume(const Phase& ph) = 0;
};

// Runs every problem to completion, all rounds executed together.
void run_problems(Batch& B, const std::vector<Prob*>& ps) {
  std::vector<Phase> phs(ps.size());
  while (true) {
    std::vector<Phase*> act;
    std::vector<Prob*> pa;
    for (size_t k = 0; k < ps.size(); ++k)
      if (!ps[k]->done) {
        phs[k] = Phase();
        ps[k]->next(phs[k]);
        act.push_back(&phs[k]);
        pa.push_back(ps[k]);
      }
    if (act.empty()) break;
    B.execute(act);
    B.pool.run((int)pa.size(), [&](int k) { pa[k]->consume(*act[k]); });
  }
}

// Lanczos eigsolve of the lowest eigenpair with thick restarts, eager convergence test and CGS2
// (KrylovKit.eigsolve(..., :SR, Lanczos) via fixedpoint, utility.jl:180-186); the same steps as
// the single-instance solver.
struct LanczosProb : Prob {
  Op op;
  long n = 0;
  double* V = nullptr;
  double* tmp = nullptr;
  const double* x0 = nullptr;
  double* xout = nullptr;
  double tol = 1e-10;
  int kmax = 30, maxiter = 300;
  bool eager = true;
  // state
  enum { S_NORM, S_START, S_EXPAND, S_RESTART, S_FINISH, S_FLIP } st = S_NORM;
  int K = 0, numiter = 1, numops = 0, converged = 0;
  double beta = 0.0, nx = 0.0, lambda = 0.0;
  std::vector<double> T, theta, U;
  double* Vc(int k) const { return V + (size_t)k * n; }
  void next(Phase& ph) override {
    const int ldt = kmax + 1;
    switch (st) {
      case S_NORM:
        ph.dot(x0, x0, n);
        break;
      case S_START:
        ph.pre.push_back(vec_item(GV_AXPBY, n, x0, Vc(0), 1.0 / nx, 0.0));
        expand_phase(ph, 0);
        break;
      case S_EXPAND:
        expand_phase(ph, K);
        break;
      case S_RESTART: {
        const int keep = (3 * kmax + 2 * converged) / 5;
        // tmp = V[:, 0..K) U[:, 0..keep): a GEMM with the coefficients uploaded
        std::vector<double> Uk((size_t)K * keep);
        for (int j = 0; j < keep; ++j)
          for (int i = 0; i < K; ++i) Uk[i + (size_t)K * j] = U[i + (size_t)K * j];
        Item g = gemm_item((int)n, keep, K, 1.0, V, 1, n, nullptr, 1, K, 0.0, tmp, n);
        g.upload = Uk;
        g.bind = 2;
        ph.pre.push_back(g);
        ph.pre.push_back(vec_item(GV_COPY, (long)keep * n, tmp, V));
        ph.pre.push_back(vec_item(GV_COPY, n, Vc(K), Vc(keep)));
        std::vector<double> f(keep);
        for (int j = 0; j < keep; ++j) f[j] = beta * U[(K - 1) + (size_t)K * j];
        std::fill(T.begin(), T.end(), 0.0);
        for (int j = 0; j < keep; ++j) {
          T[j + (size_t)ldt * j] = theta[j];
          T[keep + (size_t)ldt * j] = f[j];
          T[j + (size_t)ldt * keep] = f[j];
        }
        K = keep;
        numiter++;
        expand_phase(ph, K);
        break;
      }
      case S_FINISH: {
        std::vector<double> u0(K);
        for (int i = 0; i < K; ++i) u0[i] = U[i];
        ph.pre.push_back(maxpy_item(n, V, n, u0, 1.0, 0.0, xout));
        ph.dot(xout, x0, n);
        break;
      }
      case S_FLIP:
        ph.pre.push_back(vec_item(GV_AXPBY, n, xout, xout, -1.0, 0.0));
        break;
    }
  }
  void expand_phase(Phase& ph, int k) {
    ph.op = &op;
    ph.src = Vc(k);
    ph.dst = Vc(k + 1);
    ph.cgs = true;
    ph.V = V;
    ph.nv = k + 1;
    ph.w = Vc(k + 1);
    ph.n = n;
  }
  void take_coefs(const Phase& ph, int k) {
    const int ldt = kmax + 1;
    numops++;
    for (int i = 0; i <= k; ++i) {
      const double h = ph.coef[i] + ph.coef[64 + i];
      T[i + (size_t)ldt * k] = h;
      T[k + (size_t)ldt * i] = h;
    }
    beta = std::sqrt(std::max(0.0, ph.coef[128]));
    K = k + 1;
  }
  void decide() {
    const int ldt = kmax + 1;
    converged = 0;
    const bool check = (K == kmax || beta <= tol || eager);
    if (check) {
      std::vector<double> A((size_t)K * K);
      for (int j = 0; j < K; ++j)
        for (int i = 0; i < K; ++i) A[i + (size_t)K * j] = 0.5 * (T[i + (size_t)ldt * j] + T[j + (size_t)ldt * i]);
      theta.assign(K, 0.0);
      U.assign((size_t)K * K, 0.0);
      sym_eig(K, A.data(), theta.data(), U.data());
      if (std::fabs(beta * U[(K - 1) + (size_t)K * 0]) <= tol) {
        converged = 1;
        st = S_FINISH;
        return;
      }
    }
    if (K < kmax)
      st = S_EXPAND;
    else if (numiter == maxiter)
      st = S_FINISH;
    else
      st = S_RESTART;
  }
  void consume(const Phase& ph) override {
    switch (st) {
      case S_NORM:
        nx = std::sqrt(std::max(0.0, ph.dotv[0]));
        if (!(nx > 0) || !std::isfinite(nx)) {
          status = IMPS_E_ARG;
          done = true;
          return;
        }
        T.assign((size_t)(kmax + 1) * (kmax + 1), 0.0);
        st = S_START;
        break;
      case S_START:
        take_coefs(ph, 0);
        decide();
        break;
      case S_EXPAND:
      case S_RESTART:
        take_coefs(ph, K);
        decide();
        break;
      case S_FINISH:
        lambda = theta[0];
        if (ph.dotv[0] < 0)
          st = S_FLIP;
        else
          done = true;
        break;
      case S_FLIP:
        done = true;
        break;
    }
  }
};

// Restarted GMRES for (a0 + a1 A) x = b with an absolute tolerance (KrylovKit.linsolve via
// envs.jl:14-19); x holds the start vector on entry and the solution on exit.
struct GmresProb : Prob {
  Op op;
  long n = 0;
  double* V = nullptr;
  double* t = nullptr;
  double* r = nullptr;
  const double* b = nullptr;
  double* x = nullptr;
  double tol = 1e-10, a0 = 1.0, a1 = -1.0;
  int kmax = 30, maxiter = 300;
  enum { S_INIT, S_CYCLE, S_STEP, S_END } st = S_INIT;
  int k = 0, numiter = 0, numops = 0;
  bool conv = false;
  double beta = 0.0;
  std::vector<double> H, cs, sn, g, y;
  double* Vc(int j) const { return V + (size_t)j * n; }
  void residual_phase(Phase& ph) {
    // t = a0 x + a1 op(x), r = b - t, then |r|^2
    ph.op = &op;
    ph.src = x;
    ph.dst = t;
    ph.post.push_back(vec_item(GV_AXPBY, n, x, t, a0, a1));
    ph.post.push_back(vec_item(GV_COPY, n, b, r));
    ph.post.push_back(vec_item(GV_AXPBY, n, t, r, -1.0, 1.0));
    ph.dot(r, r, n);
  }
  void step_phase(Phase& ph, int j) {
    ph.op = &op;
    ph.src = Vc(j);
    ph.dst = Vc(j + 1);
    ph.post.push_back(vec_item(GV_AXPBY, n, Vc(j), Vc(j + 1), a0, a1));
    ph.cgs = true;
    ph.V = V;
    ph.nv = j + 1;
    ph.w = Vc(j + 1);
    ph.n = n;
  }
  void next(Phase& ph) override {
    switch (st) {
      case S_INIT:
        residual_phase(ph);
        break;
      case S_CYCLE:
        ph.pre.push_back(vec_item(GV_AXPBY, n, r, Vc(0), 1.0 / beta, 0.0));
        step_phase(ph, 0);
        break;
      case S_STEP:
        step_phase(ph, k);
        break;
      case S_END: {
        const int ldh = kmax + 1;
        y.assign(k, 0.0);
        for (int i = k - 1; i >= 0; --i) {
          double s = g[i];
          for (int j = i + 1; j < k; ++j) s -= H[i + (size_t)ldh * j] * y[j];
          y[i] = s / H[i + (size_t)ldh * i];
        }
        ph.pre.push_back(maxpy_item(n, V, n, y, 1.0, 1.0, x));
        residual_phase(ph);
        break;
      }
    }
  }
  void consume(const Phase& ph) override {
    const int ldh = kmax + 1;
    switch (st) {
      case S_INIT:
      case S_END:
        numops++;
        beta = std::sqrt(std::max(0.0, ph.dotv[0]));
        if (beta < tol) {
          conv = true;
          done = true;
          return;
        }
        if (st == S_END && numiter >= maxiter) {
          done = true;
          return;
        }
        st = S_CYCLE;
        break;
      case S_CYCLE:
      case S_STEP: {
        if (st == S_CYCLE) {
          numiter++;
          H.assign((size_t)ldh * kmax, 0.0);
          cs.assign(kmax, 0.0);
          sn.assign(kmax, 0.0);
          g.assign(kmax + 1, 0.0);
          g[0] = beta;
          k = 0;
        }
        numops++;
        for (int i = 0; i <= k; ++i) H[i + (size_t)ldh * k] = ph.coef[i] + ph.coef[64 + i];
        H[(k + 1) + (size_t)ldh * k] = std::sqrt(std::max(0.0, ph.coef[128]));
        for (int i = 0; i < k; ++i) {
          const double a = H[i + (size_t)ldh * k], bb = H[(i + 1) + (size_t)ldh * k];
          H[i + (size_t)ldh * k] = cs[i] * a + sn[i] * bb;
          H[(i + 1) + (size_t)ldh * k] = -sn[i] * a + cs[i] * bb;
        }
        const double a = H[k + (size_t)ldh * k], bb = H[(k + 1) + (size_t)ldh * k];
        const double rho = std::hypot(a, bb);
        cs[k] = a / rho;
        sn[k] = bb / rho;
        H[k + (size_t)ldh * k] = rho;
        H[(k + 1) + (size_t)ldh * k] = 0.0;
        g[k + 1] = -sn[k] * g[k];
        g[k] = cs[k] * g[k];
        k++;
        st = (std::fabs(g[k]) <= tol || k == kmax) ? S_END : S_STEP;
        break;
      }
    }
  }
};

// Dominant eigenvector by restarted Arnoldi with the eager test (the gauge fixed point of
// uniform_leftorth!/uniform_rightorth!, ortho.jl:109-202; KrylovKit.schursolve via
// gauge_fixedpoint, utility.jl:204-211).
struct ArnoldiProb : Prob {
  Op op;
  long n = 0;
  double* V = nullptr;
  const double* x0 = nullptr;
  double* xout = nullptr;
  double tol = 1e-10;
  int kmax = 30, maxiter = 300;
  enum { S_NORM, S_START, S_STEP, S_FINISH } st = S_NORM;
  int K = 0, numiter = 1, numops = 0;
  bool conv = false;
  double nx = 0.0, lam = 0.0;
  const double* start = nullptr;
  std::vector<double> H, u;
  double* Vc(int k) const { return V + (size_t)k * n; }
  void step_phase(Phase& ph, int k) {
    ph.op = &op;
    ph.src = Vc(k);
    ph.dst = Vc(k + 1);
    ph.cgs = true;
    ph.V = V;
    ph.nv = k + 1;
    ph.w = Vc(k + 1);
    ph.n = n;
  }
  void next(Phase& ph) override {
    switch (st) {
      case S_NORM:
        ph.dot(x0, x0, n);
        break;
      case S_START:
        ph.pre.push_back(vec_item(GV_AXPBY, n, start, Vc(0), 1.0 / nx, 0.0));
        step_phase(ph, 0);
        break;
      case S_STEP:
        step_phase(ph, K);
        break;
      case S_FINISH:
        ph.pre.push_back(maxpy_item(n, V, n, u, 1.0, 0.0, xout));
        if (!(conv || numiter == maxiter)) ph.dot(xout, xout, n);
        break;
    }
  }
  void consume(const Phase& ph) override {
    const int ldh = kmax + 1;
    switch (st) {
      case S_NORM:
        nx = std::sqrt(std::max(0.0, ph.dotv[0]));
        if (!(nx > 0) || !std::isfinite(nx)) {
          status = IMPS_E_ARG;
          done = true;
          return;
        }
        start = x0;
        st = S_START;
        break;
      case S_START:
      case S_STEP: {
        if (st == S_START) {
          H.assign((size_t)ldh * kmax, 0.0);
          K = 0;
        }
        numops++;
        for (int i = 0; i <= K; ++i) H[i + (size_t)ldh * K] = ph.coef[i] + ph.coef[64 + i];
        const double beta = std::sqrt(std::max(0.0, ph.coef[128]));
        H[(K + 1) + (size_t)ldh * K] = beta;
        K++;
        u.assign(K, 0.0);
        double lim = 0.0, last = 0.0;
        const int hrc = hess_dominant_c(K, H.data(), ldh, &lam, &lim, u.data(), &last);
        if (hrc == 2) {
          status = IMPS_E_SOLVER;
          done = true;
          return;
        }
        if (std::fabs(beta) * last <= tol) {
          conv = true;
          st = S_FINISH;
        } else {
          st = (K == kmax) ? S_FINISH : S_STEP;
        }
        break;
      }
      case S_FINISH:
        if (conv || numiter == maxiter) {
          done = true;
          return;
        }
        nx = std::sqrt(std::max(0.0, ph.dotv[0]));
        start = xout;
        numiter++;
        st = S_START;
        break;
    }
  }
};

}  // namespace

namespace {

constexpr double kBTolMin = 1.0e-14, kBTolMax = 1.0e-4;

double b_updatetol(double factor, int it, double eps, double tmin = kBTolMin, double tmax = kBTolMax) {
  it = std::max(it, 1);
  return std::min(std::max(eps * factor / std::sqrt((double)it), tmin), tmax);
}

// ---- batched QRpos / LQpos
struct QRJob {
  int mm, nn, trans;
  bool want_q;
  const double* in;
  long ldi;
  double* q;
  long ldq;
  double* r;
  long ldr;
  Inst* in_inst;
  int slot;  // 0: the instance's first scratch pair, 1: the second
};

void run_qr(Batch& B, const std::vector<QRJob>& jobs) {
  if (jobs.empty()) return;
  std::vector<GQ> items;
  for (const QRJob& j : jobs) {
    brequire(j.mm <= QR_MAXM && j.nn <= QR_MAXM / 2 && j.mm >= j.nn, IMPS_E_UNSUPPORTED,
             "batch: QR shape outside the batched kernel's range");
    GQ g{};
    g.mm = j.mm;
    g.nn = j.nn;
    g.trans = j.trans;
    g.want_q = j.want_q ? 1 : 0;
    g.in = j.in;
    g.ldi = j.ldi;
    g.qout = j.q;
    g.ldq = j.ldq;
    g.rout = j.r;
    g.ldr = j.ldr;
    g.W = j.slot ? j.in_inst->qW2 : j.in_inst->qW;
    g.Q = j.slot ? j.in_inst->qQ2 : j.in_inst->qQ;
    items.push_back(g);
  }
  const GQ* dq = B.ar.put(items.data(), items.size());
  const unsigned ng = (unsigned)items.size();
  int mmax = 0;
  for (const GQ& g : items) mmax = std::max(mmax, g.mm);
  hipStream_t st = B.s;
  if (mmax <= WY_MAXM)
    B.enqueue([=] { hipLaunchKernelGGL(gqr_wy_kernel, dim3(ng), dim3(QR_THREADS), 0, st, dq); });
  else if (mmax <= 1024)
    B.enqueue([=] { hipLaunchKernelGGL(gqr_kernel<16>, dim3(ng), dim3(QR_THREADS), 0, st, dq); });
  else
    B.enqueue([=] { hipLaunchKernelGGL(gqr_kernel<32>, dim3(ng), dim3(QR_THREADS), 0, st, dq); });
}

// ---- uniform_leftorth! / uniform_rightorth! (ortho.jl:109-202) for one-site cells, lockstep:
// every active instance does one QR (LQ) sweep per round; from the tenth sweep on, each sweep is
// preceded by the Arnoldi fixed point of the mixed transfer matrix (tolerance eps^2).
// left:  A raw -> AL, C (C0 from c0); right: A -> AR, C.  Returns the sweeps per instance.
std::vector<int> gauge_sweeps(Batch& B, const std::vector<Inst*>& in, bool left, const std::vector<const double*>& A,
                              const std::vector<const double*>& C0, const std::vector<double>& tol, int maxiter,
                              int eig_miniter = 10) {
  const size_t P = in.size();
  std::vector<int> it(P, 0);
  std::vector<double> eps(P, std::numeric_limits<double>::infinity());
  std::vector<char> act(P, 1);
  // Q (AL or AR) of a sweep is only read by the fixed-point step before the next sweep (from the
  // eig_miniter-th sweep on) and at the end: earlier sweeps compute R (L) alone, and an instance
  // whose last sweep skipped Q gets it from the same M once it has converged
  std::vector<char> have_q(P, 0);
  // C = C0 / |C0|
  {
    std::vector<Item> items;
    std::vector<Phase> ph(P);
    std::vector<Phase*> pp;
    for (size_t k = 0; k < P; ++k) {
      Inst& I = *in[k];
      if (C0[k] != I.C) ph[k].pre.push_back(vec_item(GV_COPY, I.nC, C0[k], I.C));
      ph[k].dot(I.C, I.C, I.nC);
      pp.push_back(&ph[k]);
    }
    B.execute(pp);
    std::vector<Phase> ph2(P);
    pp.clear();
    for (size_t k = 0; k < P; ++k) {
      Inst& I = *in[k];
      const double nrm = std::sqrt(std::max(0.0, ph[k].dotv[0]));
      ph2[k].pre.push_back(vec_item(GV_AXPBY, I.nC, I.C, I.C, nrm > 0 ? 1.0 / nrm : 1.0, 0.0));
      pp.push_back(&ph2[k]);
    }
    B.execute(pp);
  }
  double* e2 = nullptr;
  while (true) {
    std::vector<size_t> live;
    for (size_t k = 0; k < P; ++k)
      if (act[k]) live.push_back(k);
    if (live.empty()) break;
    // Arnoldi refresh of C for the instances past eig_miniter sweeps
    std::vector<ArnoldiProb> ap;
    std::vector<size_t> ak;
    for (size_t k : live)
      if (it[k] >= eig_miniter) ak.push_back(k);
    if (!ak.empty()) {
      ap.resize(ak.size());
      std::vector<Prob*> ps;
      for (size_t j = 0; j < ak.size(); ++j) {
        Inst& I = *in[ak[j]];
        ArnoldiProb& a = ap[j];
        a.op.kind = left ? OP_PUSHL : OP_PUSHR;
        a.op.in = &I;
        a.op.above = A[ak[j]];
        a.op.below = left ? I.AL : I.AR;
        a.n = (long)I.nC;
        a.V = I.VR;
        a.x0 = I.C;
        a.xout = I.ev;
        a.tol = b_updatetol(1.0, 1, eps[ak[j]] * eps[ak[j]]);
        a.kmax = (int)std::min<size_t>(30, I.nC);
        a.maxiter = 300;
        ps.push_back(&a);
      }
      run_problems(B, ps);
      std::vector<QRJob> qj;
      for (size_t j = 0; j < ak.size(); ++j) {
        Inst& I = *in[ak[j]];
        brequire(ap[j].status == IMPS_OK, ap[j].status, "batch: gauge fixed point failed");
        // left: C = R of QRpos(ev); right: C = L of LQpos(ev)
        qj.push_back(QRJob{I.D, I.D, left ? 0 : 1, false, I.ev, I.D, I.Qdump, I.D, I.C, I.D, &I, 0});
      }
      run_qr(B, qj);
    }
    // the sweep
    std::vector<Item*> s0, s1;
    std::vector<Item> c0(live.size()), g0(live.size());
    std::vector<QRJob> qj;
    e2 = B.sc.take(P);
    std::vector<GN> items;
    for (size_t j = 0; j < live.size(); ++j) {
      const size_t k = live[j];
      Inst& I = *in[k];
      const int D = I.D, d = I.d;
      c0[j] = vec_item(GV_COPY, I.nC, I.C, I.Cpre);
      s0.push_back(&c0[j]);
      const bool wq = it[k] + 1 >= eig_miniter;
      have_q[k] = wq ? 1 : 0;
      if (left) {
        // M = C A (D x dD), QR: AL = Q, C = R
        g0[j] = gemm_item(D, d * D, D, 1.0, I.C, 1, D, A[k], 1, D, 0.0, I.Mg, D);
        qj.push_back(QRJob{d * D, D, 0, wq, I.Mg, (long)d * D, I.AL, (long)d * D, I.C, D, &I, 0});
      } else {
        // M = A C (dD x D as D x dD), LQ: C = L, AR = Q
        g0[j] = gemm_item(d * D, D, D, 1.0, A[k], 1, (long)d * D, I.C, 1, D, 0.0, I.Mg, (long)d * D);
        qj.push_back(QRJob{d * D, D, 1, wq, I.Mg, D, I.AR, D, I.C, D, &I, 0});
      }
      s1.push_back(&g0[j]);
      GN g{};
      g.n = (long)I.nC;
      g.C = I.C;
      g.Cpre = I.Cpre;
      g.out = e2 + k;
      items.push_back(g);
    }
    B.run_items(s0);
    B.run_items(s1);
    run_qr(B, qj);
    const GN* dn = B.ar.put(items.data(), items.size());
    const unsigned ng = (unsigned)items.size();
    hipStream_t st = B.s;
    B.enqueue([=] { hipLaunchKernelGGL(gnormdiff_kernel, dim3(ng), dim3(256), 0, st, dn); });
    B.submit();
    BCHK(hipMemcpyAsync(B.hres, e2, P * sizeof(double), hipMemcpyDeviceToHost, B.s));
    B.rounds++;
    B.sync();
    for (size_t k : live) {
      eps[k] = std::sqrt(std::max(0.0, B.hres[k]));
      it[k]++;
      if (eps[k] < tol[k] || it[k] > maxiter) act[k] = 0;
    }
  }
  // Q of the last sweep where it was skipped (R/L recomputed into scratch; C keeps the normalised one)
  std::vector<QRJob> fin;
  for (size_t k = 0; k < P; ++k) {
    if (have_q[k]) continue;
    Inst& I = *in[k];
    const int D = I.D, d = I.d;
    if (left)
      fin.push_back(QRJob{d * D, D, 0, true, I.Mg, (long)d * D, I.AL, (long)d * D, I.Qdump, D, &I, 0});
    else
      fin.push_back(QRJob{d * D, D, 1, true, I.Mg, D, I.AR, D, I.Qdump, D, &I, 0});
  }
  run_qr(B, fin);
  return it;
}

// ---- environments: DMRGCache (envs.jl:321-518) for one-site cells, every instance in lockstep
// over the levels; identity diagonal levels are GMRES solves of x - T x + tr(rho x) I = rhs with
// the old level as start vector (recalculate!, envs.jl:530-546), other nonzero diagonals GMRES on
// the slice push, zero diagonals direct.
bool blk_zero(const Inst& I, int l, int i) {
  for (int s = 0; s < I.d; ++s)
    for (int u = 0; u < I.d; ++u)
      if (I.Wh[l + (size_t)I.w * (u + (size_t)I.d * (i + (size_t)I.w * s))] != 0.0) return false;
  return true;
}
bool blk_eye(const Inst& I, int l, int i) {
  for (int s = 0; s < I.d; ++s)
    for (int u = 0; u < I.d; ++u)
      if (std::fabs(I.Wh[l + (size_t)I.w * (u + (size_t)I.d * (i + (size_t)I.w * s))] - (u == s ? 1.0 : 0.0)) > 1e-14)
        return false;
  return true;
}
void blk_get(const Inst& I, int l, int i, double* M) {
  for (int s = 0; s < I.d; ++s)
    for (int u = 0; u < I.d; ++u)
      M[u + I.d * s] = I.Wh[l + (size_t)I.w * (u + (size_t)I.d * (i + (size_t)I.w * s))];
}

// out = sum over the terms of slice pushes of A; terms (L level or identity, block) per instance;
// Y/Z temporaries per instance; all instances' t-th terms grouped
struct STerm {
  const double* L;  // nullptr: identity
  double M[64];
};

void slice_sums(Batch& B, const std::vector<Inst*>& in, bool left, const std::vector<std::vector<STerm>>& terms,
                const std::vector<double*>& out) {
  size_t maxt = 0;
  for (auto& t : terms) maxt = std::max(maxt, t.size());
  const size_t P = in.size();
  for (size_t t = 0; t < maxt; ++t) {
    std::vector<Item> gs;
    std::vector<GM> mx;
    for (size_t k = 0; k < P; ++k) {
      if (t >= terms[k].size()) continue;
      Inst& I = *in[k];
      const int D = I.D, d = I.d;
      const double* A = left ? I.AL : I.AR;
      const STerm& st = terms[k][t];
      const double* src = A;
      if (st.L) {
        if (left)
          gs.push_back(gemm_item(D, d * D, D, 1.0, st.L, 1, (long)D * I.w, A, 1, D, 0.0, I.Ypc, D));
        else
          gs.push_back(gemm_item(d * D, D, D, 1.0, A, 1, (long)d * D, st.L, 1, (long)D * I.w, 0.0, I.Ypc, (long)d * D));
        src = I.Ypc;
      }
      GM m{};
      m.D1 = D;
      m.D2 = D;
      std::memcpy(m.M, st.M, sizeof(m.M));
      m.in = src;
      m.out = I.Zpc;
      m.beta = t == 0 ? 0.0 : 1.0;
      mx.push_back(m);
    }
    std::vector<Item*> gp;
    for (Item& g : gs) gp.push_back(&g);
    B.run_items(gp);
    if (!mx.empty()) {
      int tb = 0;
      for (GM& m : mx) {
        m.b0 = tb;
        tb += ((m.D1 + 255) / 256) * m.D2;
      }
      const GM* dm = B.ar.put(mx.data(), mx.size());
      const int ng = (int)mx.size(), dd = B.d;
      hipStream_t st = B.s;
      B.enqueue([=] { hipLaunchKernelGGL(gmix_kernel, dim3(tb), dim3(256), 0, st, dm, ng, dd); });
    }
  }
  std::vector<Item> gs;
  for (size_t k = 0; k < P; ++k) {
    Inst& I = *in[k];
    const int D = I.D, d = I.d;
    const long ldo = (long)D * I.w;
    if (terms[k].empty()) continue;  // the level stays zero (the environments were cleared)
    if (left)
      gs.push_back(gemm_item(D, D, d * D, 1.0, I.AL, (long)d * D, 1, I.Zpc, 1, (long)d * D, 0.0, out[k], ldo));
    else
      gs.push_back(gemm_item(D, D, d * D, 1.0, I.Zpc, 1, D, I.AR, D, 1, 0.0, out[k], ldo));
  }
  std::vector<Item*> gp;
  for (Item& g : gs) gp.push_back(&g);
  B.run_items(gp);
}

void envs_batch(Batch& B, const std::vector<Inst*>& in, const std::vector<double>& tol, int maxiter, int krylovdim) {
  const size_t P = in.size();
  const int w = B.w;
  // keep the old levels as GMRES start vectors, reset the environments
  {
    std::vector<Item> its;
    for (Inst* I : in) {
      if (I->envs_valid) {
        its.push_back(vec_item(GV_COPY, I->nE, I->GL, I->oldGL));
        its.push_back(vec_item(GV_COPY, I->nE, I->GR, I->oldGR));
      }
      its.push_back(vec_item(GV_FILL, I->nE, nullptr, I->GL, 0.0));
      its.push_back(vec_item(GV_FILL, I->nE, nullptr, I->GR, 0.0));
    }
    // copies before fills: two positions
    std::vector<Item*> a, b;
    for (Item& x : its) (x.v.op == GV_COPY ? a : b).push_back(&x);
    B.run_items(a);
    B.run_items(b);
    std::vector<Item> id;
    for (Inst* I : in) {
      id.push_back(ident_item(I->D, I->GL, (long)I->D * w));
      id.push_back(ident_item(I->D, I->GR + (size_t)I->D * (w - 1), (long)I->D * w));
    }
    std::vector<Item*> c;
    for (Item& x : id) c.push_back(&x);
    B.run_items(c);
  }
  // rho_L = C C^T (left levels), rho_R = C^T C (right levels)
  auto make_rho = [&](bool left) {
    std::vector<Item> g;
    for (Inst* I : in) {
      const int D = I->D;
      if (left)
        g.push_back(gemm_item(D, D, D, 1.0, I->C, 1, D, I->C, D, 1, 0.0, I->rho, D));
      else
        g.push_back(gemm_item(D, D, D, 1.0, I->C, D, 1, I->C, 1, D, 0.0, I->rho, D));
    }
    std::vector<Item*> gp;
    for (Item& x : g) gp.push_back(&x);
    B.run_items(gp);
  };
  auto solve_level = [&](bool left, int i) {
    // per instance: identity diagonal -> regularised GMRES; other nonzero diagonal -> GMRES on the
    // slice push; zero diagonal -> the right-hand side is the level
    std::vector<GmresProb> gp(P);
    std::vector<Prob*> ps;
    std::vector<size_t> ks;
    std::vector<Item> regs;
    for (size_t k = 0; k < P; ++k) {
      Inst& I = *in[k];
      const int D = I.D;
      const bool ident = (i == 0 && !left) || (i == w - 1 && left) || blk_eye(I, i, i);
      if (!ident && blk_zero(I, i, i)) continue;
      double* G = (left ? I.GL : I.GR) + (size_t)D * i;
      const double* old = (left ? I.oldGL : I.oldGR) + (size_t)D * i;
      GmresProb& g = gp[k];
      g.op.in = &I;
      if (ident) {
        g.op.kind = left ? OP_PUSHL : OP_PUSHR;
        g.op.above = left ? I.AL : I.AR;
        g.op.below = left ? I.AL : I.AR;
        g.op.rho = I.rho;
      } else {
        g.op.kind = left ? OP_SLICEL : OP_SLICER;
        g.op.above = left ? I.AL : I.AR;
        g.op.below = left ? I.AL : I.AR;
        blk_get(I, i, i, g.op.M);
      }
      g.n = (long)I.nC;
      g.V = I.VG;
      g.t = I.gt;
      g.r = I.gr;
      g.b = I.grhs;
      g.x = I.gx;
      g.tol = tol[k];
      g.kmax = (int)std::min<long>((long)I.nC, krylovdim);
      g.maxiter = maxiter;
      ps.push_back(&g);
      ks.push_back(k);
      // rhs and start vector
      regs.push_back(copy2d_item(D, D, G, (long)D * w, I.grhs, D));
      if (I.envs_valid)
        regs.push_back(copy2d_item(D, D, old, (long)D * w, I.gx, D));
      else
        regs.push_back(vec_item(GV_FILL, I.nC, nullptr, I.gx, 0.0));
    }
    if (ps.empty()) return;
    std::vector<Item*> rp;
    for (Item& x : regs) rp.push_back(&x);
    B.run_items(rp);
    run_problems(B, ps);
    std::vector<Item> back;
    std::vector<GR> rg;
    for (size_t j = 0; j < ks.size(); ++j) {
      Inst& I = *in[ks[j]];
      const int D = I.D;
      I.gmres_ops += gp[ks[j]].numops;
      double* G = (left ? I.GL : I.GR) + (size_t)D * i;
      back.push_back(copy2d_item(D, D, I.gx, D, G, (long)D * w));
      if (gp[ks[j]].op.rho) {
        GR r{};
        r.n = D;
        r.ldv = (long)D * w;
        r.v = G;
        r.l = I.rho;
        rg.push_back(r);
      }
    }
    std::vector<Item*> bp;
    for (Item& x : back) bp.push_back(&x);
    B.run_items(bp);
    if (!rg.empty()) {
      const GR* dr = B.ar.put(rg.data(), rg.size());
      const unsigned ng = (unsigned)rg.size();
      hipStream_t st = B.s;
      B.enqueue([=] { hipLaunchKernelGGL(greg_kernel, dim3(ng), dim3(1024), 0, st, dr); });
    }
  };
  auto level_empty = [&](const Inst& I, int i) {
    return i != 0 && i != w - 1 && blk_zero(I, i, i) && blk_zero(I, 0, i) && blk_zero(I, i, w - 1);
  };
  // left levels 1..w-1 (the closing level w-1 last)
  make_rho(true);
  for (int i = 1; i <= w - 1; ++i) {
    std::vector<std::vector<STerm>> terms(P);
    std::vector<double*> out(P);
    for (size_t k = 0; k < P; ++k) {
      Inst& I = *in[k];
      out[k] = I.GL + (size_t)I.D * i;
      if (i < w - 1 && level_empty(I, i)) continue;
      for (int l = 0; l < i; ++l) {
        if (blk_zero(I, l, i)) continue;
        STerm t{};
        t.L = l == 0 ? nullptr : I.GL + (size_t)I.D * l;
        blk_get(I, l, i, t.M);
        terms[k].push_back(t);
      }
    }
    slice_sums(B, in, true, terms, out);
    solve_level(true, i);
  }
  // right levels w-2..0
  make_rho(false);
  for (int i = w - 2; i >= 0; --i) {
    std::vector<std::vector<STerm>> terms(P);
    std::vector<double*> out(P);
    for (size_t k = 0; k < P; ++k) {
      Inst& I = *in[k];
      out[k] = I.GR + (size_t)I.D * i;
      if (i > 0 && level_empty(I, i)) continue;
      for (int l = i + 1; l < w; ++l) {
        if (blk_zero(I, i, l)) continue;
        STerm t{};
        t.L = l == w - 1 ? nullptr : I.GR + (size_t)I.D * l;
        blk_get(I, i, l, t.M);
        terms[k].push_back(t);
      }
    }
    slice_sums(B, in, false, terms, out);
    solve_level(false, i);
  }
  for (Inst* I : in) I->envs_valid = true;
}

}  // namespace

namespace {

void alloc_inst(Inst& I, int kd, int kde) {
  const int D = I.D, d = I.d, w = I.w, wd = w * d;
  I.nA = (size_t)D * d * D;
  I.nC = (size_t)D * D;
  I.nE = (size_t)D * w * D;
  I.kdA = (int)std::min<size_t>(kd, I.nA);
  I.kdC = (int)std::min<size_t>(kd, I.nC);
  I.kdE = (int)std::min<size_t>(kde, I.nC);
  brequire(I.kdA < kMaxBasis && I.kdC < kMaxBasis && I.kdE < kMaxBasis, IMPS_E_UNSUPPORTED,
           "batch: krylovdim must be below 32");
  const size_t nT = (size_t)D * w * d * D;
  struct Slot {
    double** p;
    size_t n;
  };
  std::vector<Slot> sl = {
      {&I.Wm, (size_t)wd * wd}, {&I.AL, I.nA}, {&I.AR, I.nA}, {&I.AC, I.nA}, {&I.pAC, I.nA}, {&I.pAL, I.nA},
      {&I.C, I.nC}, {&I.pC, I.nC}, {&I.GL, I.nE}, {&I.GR, I.nE}, {&I.oldGL, I.nE}, {&I.oldGR, I.nE},
      {&I.T1, nT}, {&I.T2, nT}, {&I.T1c, I.nE}, {&I.Ypc, I.nA}, {&I.Zpc, I.nA},
      {&I.VA, (size_t)(I.kdA + 1) * I.nA}, {&I.tmpA, (size_t)I.kdA * I.nA},
      {&I.VC, (size_t)(I.kdC + 1) * I.nC}, {&I.tmpC, (size_t)I.kdC * I.nC},
      {&I.VG, (size_t)(I.kdE + 1) * I.nC}, {&I.gt, I.nC}, {&I.gr, I.nC}, {&I.gx, I.nC}, {&I.grhs, I.nC},
      {&I.VR, (size_t)31 * I.nC}, {&I.ev, I.nC},
      {&I.qW, I.nA}, {&I.qQ, I.nA}, {&I.qW2, I.nA}, {&I.qQ2, I.nA}, {&I.Qa, I.nA}, {&I.Qc, I.nC},
      {&I.Rq, I.nC}, {&I.Rq2, I.nC}, {&I.Cpre, I.nC}, {&I.Mg, I.nA}, {&I.Qdump, I.nC}, {&I.gk, I.nA},
      {&I.gq, I.nC}, {&I.Id, I.nC}, {&I.rho, I.nC}};
  size_t tot = 0;
  for (const Slot& s : sl) tot += align_up(s.n, 32);
  BCHK(hipMalloc((void**)&I.mem, tot * sizeof(double)));
  size_t off = 0;
  for (const Slot& s : sl) {
    *s.p = I.mem + off;
    off += align_up(s.n, 32);
  }
  I.Y = I.Ypc;
  I.Z = I.Zpc;
}

// calc_galerkin (vumps.jl:111-124): |P[H_AC(AC) / |H_AC(AC)|]| per instance, two rounds
void galerkin_batch(Batch& B, const std::vector<Inst*>& in) {
  const size_t P = in.size();
  std::vector<Op> ops(P);
  std::vector<Phase> ph(P);
  std::vector<Phase*> pp;
  for (size_t k = 0; k < P; ++k) {
    Inst& I = *in[k];
    ops[k].kind = OP_HAC;
    ops[k].in = &I;
    ops[k].prec = PREC_FP64;
    ph[k].op = &ops[k];
    ph[k].src = I.AC;
    ph[k].dst = I.gk;
    ph[k].dot(I.gk, I.gk, (long)I.nA);
    pp.push_back(&ph[k]);
  }
  B.execute(pp);
  std::vector<Phase> p2(P);
  pp.clear();
  for (size_t k = 0; k < P; ++k) {
    Inst& I = *in[k];
    const int D = I.D, d = I.d;
    const double nx = std::sqrt(std::max(0.0, ph[k].dotv[0]));
    p2[k].pre.push_back(vec_item(GV_AXPBY, (long)I.nA, I.gk, I.gk, nx > 0 ? 1.0 / nx : 1.0, 0.0));
    p2[k].pre.push_back(gemm_item(D, D, d * D, 1.0, I.AL, (long)d * D, 1, I.gk, 1, (long)d * D, 0.0, I.gq, D));
    p2[k].pre.push_back(gemm_item(d * D, D, D, -1.0, I.AL, 1, (long)d * D, I.gq, 1, D, 1.0, I.gk, (long)d * D));
    p2[k].dot(I.gk, I.gk, (long)I.nA);
    pp.push_back(&p2[k]);
  }
  B.execute(pp);
  for (size_t k = 0; k < P; ++k) in[k]->eps = std::sqrt(std::max(0.0, p2[k].dotv[0]));
}

// expectationvalue(psi, H, envs) (expval.jl:93-112) / dot(psi, psi) (canonicalmps.jl:262-269)
void energy_batch(Batch& B, const std::vector<Inst*>& in) {
  const size_t P = in.size();
  std::vector<Op> ops(P);
  std::vector<Phase> ph(P);
  std::vector<Phase*> pp;
  for (size_t k = 0; k < P; ++k) {
    Inst& I = *in[k];
    ops[k].kind = OP_HACFULL;
    ops[k].in = &I;
    ph[k].op = &ops[k];
    ph[k].src = I.AC;
    ph[k].dst = nullptr;
    ph[k].dot(I.AC, I.T2 + (size_t)I.D * I.d * I.D * (I.w - 1), (long)I.nA);
    ph[k].pre.push_back(ident_item(I.D, I.Id, I.D));
    pp.push_back(&ph[k]);
  }
  B.execute(pp);
  std::vector<ArnoldiProb> ap(P);
  std::vector<Prob*> ps;
  for (size_t k = 0; k < P; ++k) {
    Inst& I = *in[k];
    ArnoldiProb& a = ap[k];
    a.op.kind = OP_PUSHL;
    a.op.in = &I;
    a.op.above = I.AL;
    a.op.below = I.AL;
    a.n = (long)I.nC;
    a.V = I.VR;
    a.x0 = I.Id;
    a.xout = I.ev;
    a.tol = 1e-10;
    a.kmax = (int)std::min<size_t>(30, I.nC);
    ps.push_back(&a);
  }
  run_problems(B, ps);
  for (size_t k = 0; k < P; ++k) {
    brequire(ap[k].status == IMPS_OK, ap[k].status, "batch: dot(psi, psi) failed");
    in[k]->f = ph[k].dotv[0] / ap[k].lam;
  }
}

// regauge! (ortho.jl:214-219): AL = Q(QRpos(AC)) Q(QRpos(C))^T
void regauge_batch(Batch& B, const std::vector<Inst*>& in) {
  std::vector<QRJob> qj;
  for (Inst* I : in) {
    const int D = I->D, d = I->d;
    qj.push_back(QRJob{d * D, D, 0, true, I->pAC, (long)d * D, I->Qa, (long)d * D, I->Rq, D, I, 0});
    qj.push_back(QRJob{D, D, 0, true, I->pC, D, I->Qc, D, I->Rq2, D, I, 1});
  }
  run_qr(B, qj);
  std::vector<Item> g;
  for (Inst* I : in) {
    const int D = I->D, d = I->d;
    g.push_back(gemm_item(d * D, D, D, 1.0, I->Qa, 1, (long)d * D, I->Qc, D, 1, 0.0, I->pAL, (long)d * D));
  }
  std::vector<Item*> gp;
  for (Item& x : g) gp.push_back(&x);
  B.run_items(gp);
}

void mul_al_c(Batch& B, const std::vector<Inst*>& in) {
  std::vector<Item> g;
  for (Inst* I : in) {
    const int D = I->D, d = I->d;
    g.push_back(gemm_item(d * D, D, D, 1.0, I->AL, 1, (long)d * D, I->C, 1, D, 0.0, I->AC, (long)d * D));
  }
  std::vector<Item*> gp;
  for (Item& x : g) gp.push_back(&x);
  B.run_items(gp);
}

// gauge_step! (vumps.jl:23-32): gaugefix!(psi, AL, C; order = :R) from the regauged AL, AC = AL C
void gauge_step_batch(Batch& B, const std::vector<Inst*>& in, const std::vector<double>& tol, int maxiter,
                      int eig_miniter) {
  std::vector<const double*> A, C0;
  for (Inst* I : in) {
    A.push_back(I->pAL);
    C0.push_back(I->C);
  }
  std::vector<int> sw = gauge_sweeps(B, in, false, A, C0, tol, maxiter, std::max(0, eig_miniter));
  std::vector<Item> cp;
  for (size_t k = 0; k < in.size(); ++k) {
    in[k]->sweeps += sw[k];
    cp.push_back(vec_item(GV_COPY, (long)in[k]->nA, in[k]->pAL, in[k]->AL));
  }
  std::vector<Item*> cpp;
  for (Item& x : cp) cpp.push_back(&x);
  B.run_items(cpp);
  mul_al_c(B, in);
}

struct BTols {
  const imps_vumps_params* p;
  double eig(int it, double eps) const {
    return p->eig_dynamic ? b_updatetol(p->eigs_tolfactor, it, eps, p->tol_min, p->tol_max) : p->eig_tol;
  }
  double gauge(int it, double eps) const {
    return p->gauge_dynamic ? b_updatetol(p->gauge_tolfactor, it, eps, p->gauge_tol_min, p->gauge_tol_max)
                            : p->gauge_tol;
  }
  double envs(int it, double eps) const {
    return p->envs_dynamic ? b_updatetol(p->envs_tolfactor, it, eps, p->envs_tol_min, p->envs_tol_max)
                           : p->envs_tol;
  }
};

void vumps_batch(Batch& B, std::vector<Inst*>& all, const imps_vumps_params* p, double gauge_tol) {
  const BTols st{p};
  auto t0 = std::chrono::steady_clock::now();
  auto secs = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  double tmark = 0.0;
  auto lap = [&](int ph) {
    B.sync();
    const double t = secs();
    B.t_phase[ph] += t - tmark;
    tmark = t;
  };
  // CanonicalIMPS(As) (canonicalmps.jl:151-170): gaugefix!(:LR) of the raw tensors (staged in pAC)
  {
    std::vector<Item> id;
    for (Inst* I : all) id.push_back(ident_item(I->D, I->Id, I->D));
    std::vector<Item*> ip;
    for (Item& x : id) ip.push_back(&x);
    B.run_items(ip);
    std::vector<const double*> A, C0;
    for (Inst* I : all) {
      A.push_back(I->pAC);
      C0.push_back(I->Id);
    }
    std::vector<double> tl(all.size(), gauge_tol);
    gauge_sweeps(B, all, true, A, C0, tl, p->gauge_maxiter);
    A.clear();
    C0.clear();
    for (Inst* I : all) {
      A.push_back(I->AL);
      C0.push_back(I->C);
    }
    gauge_sweeps(B, all, false, A, C0, tl, p->gauge_maxiter);
    mul_al_c(B, all);
    lap(0);
  }
  const long hac0 = 0;
  (void)hac0;
  for (Inst* I : all) I->hac_applies = 0;
  {
    std::vector<double> t10(all.size(), 1.0e-10);
    envs_batch(B, all, t10, p->envs_maxiter, p->envs_krylovdim);
    galerkin_batch(B, all);
    std::vector<double> te;
    for (Inst* I : all) te.push_back(st.envs(0, I->eps));
    envs_batch(B, all, te, p->envs_maxiter, p->envs_krylovdim);
    lap(3);
  }
  for (Inst* I : all) {
    I->prec = p->prec;
    I->active = true;
  }
  for (int it = 1; it <= p->maxiter; ++it) {
    std::vector<Inst*> act;
    for (Inst* I : all)
      if (I->active) act.push_back(I);
    if (act.empty()) break;
    const int siter = it - 1;
    const size_t P = act.size();
    // mixed precision: FP64 below refine_below or once the reduced phase stalls
    for (Inst* I : act)
      if (I->prec != PREC_FP64 && (I->eps < p->refine_below || (it > 4 && I->stalled))) {
        I->prec = PREC_FP64;
        I->switch_iter = it;
      }
    // localupdate_step!: AC and C eigensolves of every instance together, then regauge!
    std::vector<LanczosProb> la(P), lc(P);
    std::vector<Prob*> ps;
    for (size_t k = 0; k < P; ++k) {
      Inst& I = *act[k];
      const double te = st.eig(siter, I.eps);
      const double floor_rel =
          I.prec == PREC_FP64 ? 0.0 : (p->eig_floor > 0.0 ? p->eig_floor : (I.prec == PREC_FP32 ? 1e-6 : 1e-4));
      LanczosProb& a = la[k];
      a.op.kind = OP_HAC;
      a.op.in = &I;
      a.op.prec = I.prec;
      a.n = (long)I.nA;
      a.V = I.VA;
      a.tmp = I.tmpA;
      a.x0 = I.AC;
      a.xout = I.pAC;
      a.tol = std::max(te, floor_rel * std::max(1.0, std::fabs(I.lam_ac)));
      a.kmax = I.kdA;
      a.maxiter = p->eig_maxiter;
      a.eager = p->eig_eager != 0;
      LanczosProb& c = lc[k];
      c.op.kind = OP_HC;
      c.op.in = &I;
      c.op.prec = I.prec;
      c.n = (long)I.nC;
      c.V = I.VC;
      c.tmp = I.tmpC;
      c.x0 = I.C;
      c.xout = I.pC;
      c.tol = std::max(te, floor_rel * std::max(1.0, std::fabs(I.lam_c)));
      c.kmax = I.kdC;
      c.maxiter = p->eig_maxiter;
      c.eager = p->eig_eager != 0;
      ps.push_back(&a);
      ps.push_back(&c);
    }
    run_problems(B, ps);
    for (size_t k = 0; k < P; ++k) {
      brequire(la[k].status == IMPS_OK && lc[k].status == IMPS_OK, IMPS_E_ARG, "batch: eigensolve failed");
      act[k]->lam_ac = la[k].lambda;
      act[k]->lam_c = lc[k].lambda;
    }
    regauge_batch(B, act);
    lap(1);
    std::vector<double> tg, te;
    for (Inst* I : act) {
      tg.push_back(st.gauge(siter, I->eps));
      te.push_back(st.envs(siter, I->eps));
    }
    gauge_step_batch(B, act, tg, p->gauge_maxiter, p->gauge_eig_miniter);
    lap(2);
    envs_batch(B, act, te, p->envs_maxiter, p->envs_krylovdim);
    lap(3);
    galerkin_batch(B, act);
    lap(4);
    energy_batch(B, act);
    lap(5);
    for (Inst* I : act) {
      I->niter = it;
      I->stalled = I->eps > 0.9 * I->best;
      I->best = std::min(I->best, I->eps);
      if (I->eps <= p->tol) {
        I->converged = true;
        I->active = false;
        I->seconds = secs();
      }
    }
  }
  for (Inst* I : all)
    if (I->active) {
      I->active = false;
      I->seconds = secs();
    }
}

}  // namespace

extern "C" {

int imps_vumps_batch_device(int device, int ninst, imps_batch_instance* inst, const imps_vumps_params* p,
                            double gauge_tol, imps_batch_stats* stats) {
  if (ninst < 0 || (ninst > 0 && !inst) || !p) return IMPS_E_ARG;
  int rc = IMPS_OK;
  std::string msg;
  auto t0 = std::chrono::steady_clock::now();
  try {
    brequire(p->maxiter >= 0 && p->krylovdim >= 1 && p->eig_maxiter >= 1 && p->gauge_maxiter >= 1 &&
                 p->envs_maxiter >= 1 && p->envs_krylovdim >= 1 && p->tol >= 0,
             IMPS_E_ARG, "imps_vumps_batch_device: invalid VUMPS parameters");
    brequire(p->prec == IMPS_PREC_FP64 || p->prec == IMPS_PREC_FP32, IMPS_E_UNSUPPORTED,
             "imps_vumps_batch_device: FP64 or FP32 bulk arithmetic only");
    brequire(!p->finalize, IMPS_E_UNSUPPORTED, "imps_vumps_batch_device: no finalize callback in a batch");
    if (ninst == 0) return IMPS_OK;
    const int w = inst[0].w, d = inst[0].d;
    for (int k = 0; k < ninst; ++k) {
      const imps_batch_instance& in = inst[k];
      brequire(in.W && in.A && in.W[0] && in.A[0], IMPS_E_ARG, "imps_vumps_batch_device: null tensors");
      brequire(in.nsites == 1, IMPS_E_UNSUPPORTED, "imps_vumps_batch_device: one-site unit cells only");
      brequire(in.w == w && in.d == d, IMPS_E_UNSUPPORTED,
               "imps_vumps_batch_device: every instance needs the same MPO level count and physical dimension");
      brequire(in.D >= 16 && in.D % 16 == 0 && in.D <= QR_MAXM / 2 && (long)d * in.D <= QR_MAXM, IMPS_E_UNSUPPORTED,
               "imps_vumps_batch_device: bond dimensions must be multiples of 16 with d D <= 2048");
      const std::vector<double> Wh(in.W[0], in.W[0] + (size_t)w * d * w * d);
      for (int u = 0; u < d; ++u)
        for (int s2 = 0; s2 < d; ++s2) {
          const double e = u == s2 ? 1.0 : 0.0;
          brequire(Wh[(size_t)w * (u + (size_t)d * ((size_t)w * s2))] == e &&
                       Wh[(w - 1) + (size_t)w * (u + (size_t)d * ((w - 1) + (size_t)w * s2))] == e,
                   IMPS_E_ARG, "imps_vumps_batch_device: levels 0 and w-1 must be unit levels");
        }
    }
    brequire(w >= 2 && w * d <= 64 && d <= 8, IMPS_E_UNSUPPORTED, "imps_vumps_batch_device: w d <= 64, d <= 8");
    // The instances are dealt, largest D^3 first, to a few lockstep groups that run concurrently
    // (one host thread and stream each): a group's host-side work and its slowest instances overlap
    // with the other groups' kernels.
    std::vector<int> order(ninst);
    for (int k = 0; k < ninst; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return (double)inst[a].D * inst[a].D * inst[a].D > (double)inst[b].D * inst[b].D * inst[b].D;
    });
    const int G = ninst >= 16 ? 4 : (ninst >= 4 ? 2 : 1);
    std::vector<std::vector<int>> members(G);
    for (int k = 0; k < ninst; ++k) members[k % G].push_back(order[k]);
    std::vector<int> grc(G, IMPS_OK);
    std::vector<std::string> gmsg(G);
    std::vector<long> grounds(G, 0), glaunch(G, 0);
    std::vector<std::vector<double>> gphase(G, std::vector<double>(6, 0.0));
    auto run_group = [&](int g) {
      Batch B;
      try {
        B.init(device, std::max(1, 8 / G - 1));
        B.w = w;
        B.d = d;
        for (int k : members[g]) {
          Inst* I = new Inst();
          B.insts.push_back(I);
          I->idx = k;
          I->D = inst[k].D;
          I->d = d;
          I->w = w;
          I->Wh.assign(inst[k].W[0], inst[k].W[0] + (size_t)w * d * w * d);
          alloc_inst(*I, p->krylovdim, p->envs_krylovdim);
          const int wd = w * d;
          std::vector<double> Wm((size_t)wd * wd);
          for (int i = 0; i < w; ++i)
            for (int u = 0; u < d; ++u)
              for (int j = 0; j < w; ++j)
                for (int s2 = 0; s2 < d; ++s2)
                  Wm[(u + (size_t)d * j) + (size_t)wd * (i + (size_t)w * s2)] =
                      I->Wh[i + (size_t)w * (u + (size_t)d * (j + (size_t)w * s2))];
          BCHK(hipMemcpyAsync(I->Wm, Wm.data(), Wm.size() * sizeof(double), hipMemcpyHostToDevice, B.s));
          BCHK(hipMemcpyAsync(I->pAC, inst[k].A[0], I->nA * sizeof(double), hipMemcpyHostToDevice, B.s));
          BCHK(hipStreamSynchronize(B.s));
        }
        vumps_batch(B, B.insts, p, gauge_tol);
        for (const Inst* I : B.insts) {
          imps_batch_instance& o = inst[I->idx];
          o.status = IMPS_OK;
          o.niter = I->niter;
          o.converged = I->converged ? 1 : 0;
          o.itererr = I->eps;
          o.energy = I->f;
          o.seconds = I->seconds;
          o.hac_applies = (int)I->hac_applies;
        }
      } catch (const BatchError& e) {
        grc[g] = e.code;
        gmsg[g] = e.what();
      } catch (const std::exception& e) {
        grc[g] = IMPS_E_ARG;
        gmsg[g] = e.what();
      }
      grounds[g] = B.rounds;
      glaunch[g] = B.launches;
      for (int k = 0; k < 6; ++k) gphase[g][k] = B.t_phase[k];
      B.release();
    };
    std::vector<std::thread> th;
    for (int g = 1; g < G; ++g) th.emplace_back(run_group, g);
    run_group(0);
    for (auto& t : th) t.join();
    for (int g = 0; g < G; ++g)
      if (grc[g] != IMPS_OK) {
        rc = grc[g];
        msg = gmsg[g];
        for (int k : members[g]) inst[k].status = rc;
      }
    if (stats) {
      stats->rounds = 0;
      stats->launches = 0;
      for (int k = 0; k < 6; ++k) stats->t_phase[k] = 0.0;
      for (int g = 0; g < G; ++g) {
        stats->rounds += grounds[g];
        stats->launches += glaunch[g];
        for (int k = 0; k < 6; ++k) stats->t_phase[k] = std::max(stats->t_phase[k], gphase[g][k]);
      }
      stats->groups = G;
      stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
  } catch (const BatchError& e) {
    rc = e.code;
    msg = e.what();
  } catch (const std::exception& e) {
    rc = IMPS_E_ARG;
    msg = e.what();
  }
  if (rc != IMPS_OK) {
    set_global_error(msg);
    for (int k = 0; k < ninst; ++k)
      if (inst[k].status == IMPS_OK) inst[k].status = rc;
  }
  return rc;
}

}  // extern "C"
