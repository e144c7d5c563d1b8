// Internal declarations shared by the libimps translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

typedef struct imps_ctx imps_ctx;

namespace imps {

// Strided GEMM on column-major data:
//   C[m + n*ldc] = alpha * sum_k opA(m,k) * opB(k,n) + beta * C[m + n*ldc]
// with k = kc*Kc + kk (0 <= kk < Kc) and
//   opA(m,k) = A[m*sAm + kk*sAk + kc*sAc],  opB(k,n) = B[kk*sBk + kc*sBc + n*sBn].
// Either sAm or sAk must be 1, and either sBk or sBn must be 1.  When Kc < K, Kc must be
// a multiple of 16.
struct GemmDesc {
  int M = 0, N = 0, K = 0, Kc = 0;
  const double* A = nullptr;
  long sAm = 1, sAk = 0, sAc = 0;
  const double* B = nullptr;
  long sBk = 1, sBn = 0, sBc = 0;
  double* C = nullptr;
  long ldc = 0;
  double alpha = 1.0, beta = 0.0;
  // optional split-K workspace of gemm_workspace(M, N, K) doubles; nullptr disables split-K
  double* W = nullptr;
};

size_t gemm_workspace(int M, int N, int K);

enum Prec { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2 };

hipError_t launch_gemm(const GemmDesc& g, int prec, hipStream_t s);
// Bit-identical tile variants of the FP64/FP32 GEMM (no split-K): fills ids with the variants
// valid for g (at most 8) and returns their number, or 0 when launch_gemm would split K for this shape (or
// prec is split-BF16).  launch_gemm_variant runs one of them.
int gemm_tile_variants(const GemmDesc& g, int prec, int* ids);
hipError_t launch_gemm_variant(const GemmDesc& g, int prec, int variant, hipStream_t s);

// out[a,u,b] = sum_s M[u,s] in[a,s,b] + beta*out[a,u,b]; in/out (D1, d, D2), M (d, d) on host.
hipError_t launch_mix_mid(int D1, int d, int D2, const double* M, const double* in, double* out,
                          double beta, hipStream_t s);

// H_AC W-mix: T2[b,u,bb,j] = sum_{i,s} W[i,u,j,s] T1[b,i,s,bb]
// T1 at b + D*(i + w*(s + d*bb)), T2 at b + D*(u + d*(bb + D*j)); Wm (w*d)x(w*d) on device
// as Wm[(u + d*j) + wd*(i + w*s)].  X0 (optional): level i = 0 read from x at b + D*(s + d*bb)
// instead of T1; Y (optional): channel j = w-1 written to y at b + D*(u + d*bb) instead of T2.
hipError_t launch_wmix(int D, int w, int d, const double* Wm_dev, const double* T1, double* T2, const double* X0,
                       double* Y, hipStream_t s);

// Split-BF16 planes (x_hi = bf16(float(x)), x_lo = bf16(float(x) - x_hi), stored as raw bf16):
// out[r*ldo + c] = split(in[r*ldi + c]) (trans = false; cols % 8 == 0) or split(in[r + ldi*c])
// (trans = true; rows % 64 == cols % 64 == 0).
hipError_t launch_split_planes(int rows, int cols, const double* in, long ldi, bool trans, uint16_t* hi,
                               uint16_t* lo, long ldo, hipStream_t s);
// C(m, n) = alpha sum_k A(m, k) B(k, n) + beta C(m, n) from planes: A(m, k) at m*lda + k,
// B(k, n) at n*ldb + k, C(m, n) at m + n*ldc (FP64 when c_double, else FP32 with beta = 0).
// Three bf16 products per element pair, FP32 accumulation.  M % 128 == 0, N % 64 == 0,
// K % 64 == 0, lda, ldb multiples of 8 and 16-byte aligned planes.
hipError_t launch_gemm_planes(int M, int N, int K, const uint16_t* Ah, const uint16_t* Al, long lda,
                              const uint16_t* Bh, const uint16_t* Bl, long ldb, void* C, long ldc, bool c_double,
                              double alpha, double beta, hipStream_t s);
// For each level z < nlev: out[z*out_z + r*ldo + c] = split(in[z*in_z + r + ldi*c]) from FP32
// (rows % 64 == cols % 64 == 0).
hipError_t launch_split_planes_levels(int rows, int cols, int nlev, const float* in, long ldi, long in_z,
                                      uint16_t* hi, uint16_t* lo, long ldo, long out_z, hipStream_t s);
// Split-K plane GEMM: S partial products over K/S each, P[z*M*N + m + M*n] (FP32);
// K % (64 S) == 0, M % 128 == 0, N % 64 == 0.
hipError_t launch_gemm_planes_splitk(int M, int N, int K, int S, const uint16_t* Ah, const uint16_t* Al, long lda,
                                     const uint16_t* Bh, const uint16_t* Bl, long ldb, float* P, hipStream_t s);
// y[a + D*b] = sum_{z<S} P[z*pstride + a + D*b] + add[a + add_ld*b] (D x D, FP64 out)
hipError_t launch_planes_reduce(int D, int S, const float* P, long pstride, const float* add, long add_ld, double* y,
                                hipStream_t s);
// The H_AC mix of launch_wmix for identity structural levels on planes: T1 in FP32 (levels
// 1..w-1 at b + D*(i + w*(s + d*bb))), level 0 from X0; channel j = w-1 to Y (FP64), channels
// j < w-1 as planes Th/Tl[(b + D*u)*D*(w-1) + D*j + bb].  D % 32 == 0, 2 <= w <= 8, d <= 4.
hipError_t launch_wmix_planes(int D, int w, int d, const double* Wm, const float* T1, const double* X0,
                              uint16_t* Th, uint16_t* Tl, double* Y, hipStream_t s);

// Energy-gradient sandwich partials: Gp[o*D + bb] = sum_b T1[b,i,s,bb] T3[b,u,bb,j] with
// o = (i + w s) + w d (u + d j); T1 as in launch_wmix, T3 laid out like its T2; w d <= 32.
hipError_t launch_gsand(int D, int w, int d, const double* T1, const double* T3, double* Gp, hipStream_t s);

// Partial sums for dots: part[v*nb + blk] for v < nv over vectors X + v*ldx against y.
int dot_blocks(size_t n);
hipError_t launch_multidot(size_t n, int nv, const double* X, size_t ldx, const double* y,
                           double* part, hipStream_t s);
// fused Gram-Schmidt sweep (nv <= 32, coef on the device): w -= X coef, then part receives the
// per-block shares of X_v . w (dots) or of w . w (!dots), laid out as launch_multidot's
hipError_t launch_axpy_reduce(size_t n, int nv, const double* X, size_t ldx, const double* coef, double* w,
                              double* part, bool dots, hipStream_t s);
// MINRES on the Jacobi-Davidson correction equation with its scalars on the device: S holds
// MR_NS doubles (the recurrence state and the coefficients of the next vector update).
enum {
  MR_DELTA = 0, MR_GAMMA, MR_ZZ, MR_BETA, MR_COLD, MR_C, MR_SOLD, MR_S, MR_ETA,
  MR_KV, MR_KWOLD, MR_KW, MR_KT, MR_INVB, MR_DONE, MR_NS
};
// part[0..nb) <- block shares of a . y, part[nb..2nb) <- of b . y (nb = dot_blocks(n))
hipError_t launch_dot2(size_t n, const double* a, const double* b, const double* y, double* part, hipStream_t s);
// z = y - S[GAMMA] x - S[DELTA] v - S[BETA] vold; part <- block shares of z . z
hipError_t launch_minres_z(size_t n, const double* y, const double* x, const double* v, const double* vold,
                           const double* S, double* z, double* part, hipStream_t s);
// one MINRES Givens step from alpha = S[DELTA] - theta and beta = sqrt(S[ZZ])
hipError_t launch_minres_scalar(double* S, double theta, hipStream_t s);
// wold <- K_V v + K_WOLD wold + K_W w; t += K_T wold; z *= INV_BETA
hipError_t launch_minres_update(size_t n, const double* v, double* wold, const double* w, double* t, double* z,
                                const double* S, hipStream_t s);
hipError_t launch_minres_init(double* S, double beta1, hipStream_t s);
// y = alpha*y + sum_v coef[v] * X[v]  (nv <= 64 coefficients passed by value)
hipError_t launch_multiaxpy(size_t n, int nv, const double* X, size_t ldx, const double* coef,
                            double alpha, double* y, hipStream_t s);
hipError_t launch_scale(size_t n, double a, double* x, hipStream_t s);
// out[v] = sum_b part[v*nb + b] (one thread per v, fixed order)
hipError_t launch_partials_reduce(int nv, int nb, const double* part, double* out, hipStream_t s);
// y = alpha*y + scale * sum_v coef[v] * X[v] with the coefficients in device memory
hipError_t launch_multiaxpy_dev(size_t n, int nv, const double* X, size_t ldx, const double* coef,
                                double scale, double alpha, double* y, hipStream_t s);
// x *= 1/sqrt(*sq) when *sq > 0 (sq in device memory)
hipError_t launch_scale_inv_sqrt(size_t n, double* x, const double* sq, hipStream_t s);
// y = a*x + b*y
hipError_t launch_axpby(size_t n, double a, const double* x, double b, double* y, hipStream_t s);
hipError_t launch_fill(size_t n, double v, double* x, hipStream_t s);
// Complex tensors (D1, m, D2), re/im interleaved, and their real embedding (2 D1, m, 2 D2) with
// each entry z as the block [[Re z, -Im z], [Im z, Re z]]: R = scale * embed(Z); Z = scale *
// (nearest complex tensor of R), R projected onto the embedded form when project (Z may be null);
// y += alpha * i w for embedded tensors of n doubles with an even leading dimension.
hipError_t launch_cplx_embed(int D1, int m, int D2, const double* Z, double scale, double* R, hipStream_t s);
hipError_t launch_cplx_extract(int D1, int m, int D2, double* R, double scale, double* Z, bool project,
                               hipStream_t s);
hipError_t launch_cplx_axpy_i(size_t n, double alpha, const double* w, double* y, hipStream_t s);
// Bond-dependent bond dimensions (active blocks inside the largest): G[i][i] = 1 for
// nact <= i < n; A[i + ld*j] = 0 for i >= ract or j >= cact.
hipError_t launch_pad_unit_diag(int n, int nact, double* G, long ld, hipStream_t s);
hipError_t launch_mask_block(int rows, int cols, int ract, int cact, double* A, long ld, hipStream_t s);
// diagonal of an n x n matrix (leading dimension ld) += v
hipError_t launch_add_diag(int n, double v, double* x, long ld, hipStream_t s);
// x[i + ld*j] = v for i < rows, j < cols
hipError_t launch_fill2d(int rows, int cols, double v, double* x, long ld, hipStream_t s);
hipError_t launch_set_identity(int n, double* x, long ld, hipStream_t s);
// out (cols x rows) = in^T, in (rows x cols) column-major with leading dims
hipError_t launch_transpose(int rows, int cols, const double* in, long ldin, double* out,
                            long ldout, hipStream_t s);
// partial sums of sum_{a,b} l[a + ldl*b] * v[b + ldv*a] (n x n), one per 32 x 32 tile
hipError_t launch_trace_prod_parts(int n, const double* l, long ldl, const double* v, long ldv, double* part,
                                   int* nparts, hipStream_t s);
// c = sum_{b < np} part[b] on the device (fixed order), then x[i + ld*i] += factor * c for i < n
// and *out = c when out != nullptr (one workgroup)
hipError_t launch_reduce_add_diag(int np, const double* part, double factor, double* x, long ld, int n, double* out,
                                  hipStream_t s);
// c = sum_{b < np} part[b] (fixed order), then v[i + ldv*j] += factor * c * fix[i + ldf*j] (n x n)
hipError_t launch_reduce_axpy_mat(int np, const double* part, double factor, const double* fix, long ldf, double* v,
                                  long ldv, int n, hipStream_t s);
// column/row sign fix for QRpos: Q[:,i] *= s_i, R[i,:] *= s_i with s_i = sign(R[i,i]);
// sbuf: n doubles of device scratch
hipError_t launch_qr_signfix(int m, int n, double* Q, long ldq, double* R, long ldr, double* sbuf,
                             hipStream_t s);
// Blocked Cholesky helper (lower factor G = L L^T, 64 x 64 diagonal blocks): factor the
// nb x nb diagonal block at A (leading dimension lda) in place, write U = L^{-T} to Ut
// (64 x 64, ld 64, zero outside the upper nb x nb triangle); *fail = 1 on a non-positive
// pivot
hipError_t launch_chol_diag(int nb, double* A, long lda, double* Ut, int* fail, hipStream_t s);
// G[i + ld i] += factor * trace(G) (n x n), the shift computed on the device
hipError_t launch_add_diag_trace(int n, double factor, double* G, long ld, hipStream_t s);
// A[j + ld i] = A[i + ld j] for i > j (mirror the lower triangle into the upper one)
hipError_t launch_mirror_lower(int n, double* A, long ld, hipStream_t s);
hipError_t launch_fill_int(int n, int v, int* x, hipStream_t s);
// zero the strictly lower (keep_upper) or strictly upper triangle of an n x n matrix
hipError_t launch_tri_keep(int n, double* A, long ld, bool keep_upper, hipStream_t s);
// copy upper triangle of the n x n leading block of A (lda) into R (n x n), zero below
hipError_t launch_copy_upper(int n, const double* A, long lda, double* R, hipStream_t s);
// strided 2-D copy: dst[i + j*ldd] = src[i + j*lds]
hipError_t launch_copy2d(int rows, int cols, const double* src, long lds, double* dst, long ldd,
                         hipStream_t s);
// deterministic pseudo-random fill in [-1, 1) (for measurement inputs)
hipError_t launch_fill_random(size_t n, uint64_t seed, double* x, hipStream_t s);

// message returned by imps_last_error(NULL) (calls without a context)
void set_global_error(const std::string& msg);
// message returned by imps_last_error(c)
void set_context_error(imps_ctx* c, const std::string& msg);
// unit cell and physical dimension of the context's model; false when no model is set
bool context_model(const imps_ctx* c, int* nsites, int* d);

// ---- host small dense linear algebra (imps_small.cpp) ----
// Symmetric eigen-decomposition of the n x n column-major A (destroyed): eigenvalues
// ascending in w, orthonormal eigenvectors in the columns of V.
void sym_eig(int n, double* A, double* w, double* V);
// Eigenvalues (wr + i wi) of an upper-Hessenberg n x n column-major matrix H (copied).
// Returns 0 on success.
int hess_eigvals(int n, const double* H, int ldh, double* wr, double* wi);
// Dominant (largest modulus) eigenpair of the n x n upper-Hessenberg H.  For a complex
// dominant pair the real part of the phase-fixed eigenvector is returned and *lam its real part.
// u is unit norm.  Returns 0 on success, 1 if the dominant value is complex.
int hess_dominant(int n, const double* H, int ldh, double* lam, double* u);
// The same with the eigenvalue's imaginary part (lam_im, may be NULL) and |x_{n-1}| of the unit
// complex eigenvector x (last, may be NULL; the Arnoldi residual is |beta| times it).
int hess_dominant_c(int n, const double* H, int ldh, double* lam_re, double* lam_im, double* u, double* last);
// The same for a complex upper Hessenberg matrix (column-major, ld ldh): eigenvalues (wr, wi),
// and the dominant eigenpair with the unit eigenvector x and |x_{n-1}|; 0 on success.
int hess_eigvals_z(int n, const std::complex<double>* H, int ldh, double* wr, double* wi);
int hess_dominant_z(int n, const std::complex<double>* H, int ldh, std::complex<double>* lam,
                    std::complex<double>* x, double* last);

}  // namespace imps
