// Host-side dense linear algebra on the small projected matrices of the Krylov
// solvers (dimension <= krylovdim + 1).  Written from the textbook algorithms:
//  - cyclic Jacobi for the symmetric (Lanczos) projection (Golub & Van Loan 8.5);
//  - Wilkinson-shifted complex QR iteration for the eigenvalues of the upper-Hessenberg
//    (Arnoldi) projection, then inverse iteration for the dominant eigenvector.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "imps_internal.h"

namespace imps {

void sym_eig(int n, double* A, double* w, double* V) {
  // A column-major n x n, symmetric (only A itself is used; it is symmetrised first)
  auto a = [&](int i, int j) -> double& { return A[i + (size_t)n * j]; };
  auto v = [&](int i, int j) -> double& { return V[i + (size_t)n * j]; };
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      const double s = 0.5 * (a(i, j) + a(j, i));
      a(i, j) = s;
      a(j, i) = s;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) v(i, j) = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += a(i, j) * a(i, j);
        if (i != j) off += a(i, j) * a(i, j);
      }
    if (off <= 1e-34 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        const double app = a(p, p), aqq = a(q, q);
        // rotation annihilating a(p,q): tau = (aqq - app) / (2 apq), t = sign/(|tau|+sqrt(1+tau^2))
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J (columns p, q)
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, q)
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        a(p, q) = 0.0;
        a(q, p) = 0.0;
        for (int k = 0; k < n; ++k) {  // V <- V J
          const double vkp = v(k, p), vkq = v(k, q);
          v(k, p) = c * vkp - s * vkq;
          v(k, q) = s * vkp + c * vkq;
        }
      }
    }
  }
  // sort ascending
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = a(i, i);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  std::vector<double> Vs((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    w[k] = d[idx[k]];
    for (int i = 0; i < n; ++i) Vs[i + (size_t)n * k] = v(i, idx[k]);
  }
  std::copy(Vs.begin(), Vs.end(), V);
}

typedef std::complex<double> cplx;

// Eigenvalues of the complex n x n matrix `a` (upper Hessenberg, column-major, overwritten) by
// the shifted complex QR iteration; hnorm is the largest entry modulus.
static int hqr_complex(int n, std::vector<cplx>& a, double hnorm, double* wr, double* wi) {
  auto A = [&](int i, int j) -> cplx& { return a[i + (size_t)n * j]; };
  if (hnorm == 0.0) hnorm = 1.0;
  const double eps = 2.220446049250313e-16;
  int hi = n - 1;
  int iter = 0;
  while (hi >= 0) {
    if (hi == 0) {
      wr[0] = A(0, 0).real();
      wi[0] = A(0, 0).imag();
      break;
    }
    // find the lowest negligible subdiagonal in the active window
    int lo = hi;
    while (lo > 0) {
      const double sub = std::abs(A(lo, lo - 1));
      const double dia = std::abs(A(lo, lo)) + std::abs(A(lo - 1, lo - 1));
      bool negligible = sub <= eps * (dia > 0.0 ? dia : hnorm);
      if (!negligible && iter >= 30) {
        // after 30 steps on one window also the criterion of Ahues and Tisseur (as in LAPACK's
        // xLAHQR), which deflates a nearly defective pair of close eigenvalues: the Hessenberg
        // matrices of Krylov spaces of operators with paired spectra (complex states in their
        // real embedding) hold such pairs, where the plain test stalls
        const double up = std::abs(A(lo - 1, lo));
        const double ab = std::max(sub, up), ba = std::min(sub, up);
        const double d1 = std::abs(A(lo, lo)), d2 = std::abs(A(lo - 1, lo - 1) - A(lo, lo));
        const double aa = std::max(d1, d2), bb = std::min(d1, d2);
        const double s = aa + ab;
        negligible = s > 0.0 && ba * (ab / s) <= eps * (bb * (aa / s));
      }
      if (negligible) {
        A(lo, lo - 1) = 0.0;
        break;
      }
      --lo;
    }
    if (lo == hi) {  // deflate one eigenvalue
      wr[hi] = A(hi, hi).real();
      wi[hi] = A(hi, hi).imag();
      --hi;
      iter = 0;
      continue;
    }
    if (lo == hi - 1 && iter >= 30) {
      // an isolated 2 x 2 block that the shifted steps do not split (a nearly defective pair, whose
      // subdiagonal stalls near sqrt(eps)): its two eigenvalues in closed form
      const cplx a11 = A(hi - 1, hi - 1), a12 = A(hi - 1, hi), a21 = A(hi, hi - 1), a22 = A(hi, hi);
      const cplx m = 0.5 * (a11 + a22), q = 0.5 * (a11 - a22);
      const cplx r = std::sqrt(q * q + a12 * a21);
      const cplx l1 = m + r, l2 = m - r;
      wr[hi] = l2.real();
      wi[hi] = l2.imag();
      wr[hi - 1] = l1.real();
      wi[hi - 1] = l1.imag();
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++iter > 300) return 1;
    // Wilkinson shift: eigenvalue of the trailing 2x2 closest to A(hi,hi)
    const cplx a11 = A(hi - 1, hi - 1), a12 = A(hi - 1, hi), a21 = A(hi, hi - 1), a22 = A(hi, hi);
    const cplx tr = a11 + a22, det = a11 * a22 - a12 * a21;
    const cplx disc = std::sqrt(tr * tr - 4.0 * det);
    const cplx l1 = 0.5 * (tr + disc), l2 = 0.5 * (tr - disc);
    cplx mu = (std::abs(l1 - a22) < std::abs(l2 - a22)) ? l1 : l2;
    if (iter % 11 == 10) mu = a22 + cplx(std::abs(a21), 0.0);  // exceptional shift
    // one QR step on the window [lo, hi] of A - mu I using Givens rotations
    std::vector<cplx> cs(hi - lo), sn(hi - lo);
    for (int k = lo; k <= hi; ++k) A(k, k) -= mu;
    for (int k = lo; k < hi; ++k) {
      const cplx x = A(k, k), y = A(k + 1, k);
      const double r = std::sqrt(std::norm(x) + std::norm(y));
      cplx c, s;
      if (r == 0.0) {
        c = 1.0;
        s = 0.0;
      } else {
        c = x / r;
        s = y / r;
      }
      cs[k - lo] = c;
      sn[k - lo] = s;
      // G^H applied to rows k, k+1: [conj(c) conj(s); -s c]
      for (int j = k; j < n; ++j) {
        const cplx u = A(k, j), v = A(k + 1, j);
        A(k, j) = std::conj(c) * u + std::conj(s) * v;
        A(k + 1, j) = -s * u + c * v;
      }
    }
    for (int k = lo; k < hi; ++k) {  // R G on columns k, k+1
      const cplx c = cs[k - lo], s = sn[k - lo];
      for (int i = 0; i <= std::min(k + 1, hi); ++i) {
        const cplx u = A(i, k), v = A(i, k + 1);
        A(i, k) = u * c + v * s;
        A(i, k + 1) = -u * std::conj(s) + v * std::conj(c);
      }
    }
    for (int k = lo; k <= hi; ++k) A(k, k) += mu;
  }
  return 0;
}

int hess_eigvals(int n, const double* H, int ldh, double* wr, double* wi) {
  std::vector<cplx> a((size_t)n * n);
  double hnorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      a[i + (size_t)n * j] = (i <= j + 1) ? cplx(H[i + (size_t)ldh * j], 0.0) : cplx(0.0, 0.0);
      hnorm = std::max(hnorm, std::abs(a[i + (size_t)n * j]));
    }
  return hqr_complex(n, a, hnorm, wr, wi);
}

int hess_eigvals_z(int n, const cplx* H, int ldh, double* wr, double* wi) {
  std::vector<cplx> a((size_t)n * n);
  double hnorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      a[i + (size_t)n * j] = (i <= j + 1) ? H[i + (size_t)ldh * j] : cplx(0.0, 0.0);
      hnorm = std::max(hnorm, std::abs(a[i + (size_t)n * j]));
    }
  return hqr_complex(n, a, hnorm, wr, wi);
}

// Solve (H - lam I) x = b by Gaussian elimination with partial pivoting (complex); H is the
// n x n upper Hessenberg matrix given entrywise by h(i, j).
template <typename HF>
static void shifted_solve_f(int n, const HF& h, cplx lam, std::vector<cplx>& x, double hnorm) {
  std::vector<cplx> M((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i)
      M[i + (size_t)n * j] = ((i <= j + 1) ? h(i, j) : cplx(0.0, 0.0)) - (i == j ? lam : cplx(0.0, 0.0));
  std::vector<cplx> b = x;
  const double tiny = 1e-14 * (hnorm > 0.0 ? hnorm : 1.0);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(M[i + (size_t)n * k]) > std::abs(M[p + (size_t)n * k])) p = i;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(M[k + (size_t)n * j], M[p + (size_t)n * j]);
      std::swap(b[k], b[p]);
    }
    if (std::abs(M[k + (size_t)n * k]) < tiny) M[k + (size_t)n * k] = tiny;
    for (int i = k + 1; i < n; ++i) {
      const cplx f = M[i + (size_t)n * k] / M[k + (size_t)n * k];
      if (f == cplx(0.0, 0.0)) continue;
      for (int j = k; j < n; ++j) M[i + (size_t)n * j] -= f * M[k + (size_t)n * j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    cplx s = b[i];
    for (int j = i + 1; j < n; ++j) s -= M[i + (size_t)n * j] * x[j];
    x[i] = s / M[i + (size_t)n * i];
  }
}

static void shifted_solve(int n, const double* H, int ldh, cplx lam, std::vector<cplx>& x, double hnorm) {
  shifted_solve_f(n, [&](int i, int j) { return cplx(H[i + (size_t)ldh * j], 0.0); }, lam, x, hnorm);
}

// Dominant (largest modulus) eigenpair of a complex upper Hessenberg matrix (the projection of
// an Arnoldi process in complex arithmetic): returns 0, the eigenvalue in *lam, the unit
// eigenvector in x (phase: largest entry real positive) and |x_{n-1}| in *last; 2 when the
// eigenvalue iteration fails.
int hess_dominant_z(int n, const cplx* H, int ldh, cplx* lam, cplx* x, double* last) {
  std::vector<double> wr(n), wi(n);
  if (hess_eigvals_z(n, H, ldh, wr.data(), wi.data()) != 0) return 2;
  int best = 0;
  for (int i = 1; i < n; ++i)
    if (std::hypot(wr[i], wi[i]) > std::hypot(wr[best], wi[best])) best = i;
  *lam = cplx(wr[best], wi[best]);
  double hnorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= std::min(j + 1, n - 1); ++i) hnorm = std::max(hnorm, std::abs(H[i + (size_t)ldh * j]));
  std::vector<cplx> v(n, cplx(1.0, 0.0));
  for (int it = 0; it < 3; ++it) {
    shifted_solve_f(n, [&](int i, int j) { return H[i + (size_t)ldh * j]; }, *lam, v, hnorm);
    double nr = 0.0;
    for (int i = 0; i < n; ++i) nr += std::norm(v[i]);
    nr = std::sqrt(nr);
    for (int i = 0; i < n; ++i) v[i] /= nr;
  }
  int k = 0;
  for (int i = 1; i < n; ++i)
    if (std::abs(v[i]) > std::abs(v[k])) k = i;
  const cplx ph = std::abs(v[k]) / v[k];
  for (int i = 0; i < n; ++i) x[i] = v[i] * ph;
  if (last) *last = std::abs(x[n - 1]);
  return 0;
}

int hess_dominant_c(int n, const double* H, int ldh, double* lam_re, double* lam_im, double* u, double* last) {
  std::vector<double> wr(n), wi(n);
  if (hess_eigvals(n, H, ldh, wr.data(), wi.data()) != 0) return 2;
  int best = 0;
  for (int i = 1; i < n; ++i)
    if (std::hypot(wr[i], wi[i]) > std::hypot(wr[best], wi[best])) best = i;
  const cplx lam(wr[best], wi[best]);
  double hnorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= std::min(j + 1, n - 1); ++i) hnorm = std::max(hnorm, std::fabs(H[i + (size_t)ldh * j]));
  std::vector<cplx> x(n, cplx(1.0, 0.0));
  for (int it = 0; it < 3; ++it) {
    std::vector<cplx> y(n);
    for (int i = 0; i < n; ++i) y[i] = x[i];
    x = y;
    shifted_solve(n, H, ldh, lam, x, hnorm);
    double nr = 0.0;
    for (int i = 0; i < n; ++i) nr += std::norm(x[i]);
    nr = std::sqrt(nr);
    for (int i = 0; i < n; ++i) x[i] /= nr;
  }
  // |x_{n-1}| of the unit (complex) eigenvector: the Arnoldi residual is |beta| times this
  if (last) *last = std::abs(x[n - 1]);
  int k = 0;
  for (int i = 1; i < n; ++i)
    if (std::abs(x[i]) > std::abs(x[k])) k = i;
  const cplx ph = std::abs(x[k]) / x[k];
  double nr = 0.0;
  for (int i = 0; i < n; ++i) {
    u[i] = (x[i] * ph).real();
    nr += u[i] * u[i];
  }
  nr = std::sqrt(nr);
  for (int i = 0; i < n; ++i) u[i] /= nr;
  *lam_re = lam.real();
  if (lam_im) *lam_im = lam.imag();
  return (std::fabs(lam.imag()) > 1e-12 * std::max(1.0, std::abs(lam))) ? 1 : 0;
}

int hess_dominant(int n, const double* H, int ldh, double* lam_out, double* u) {
  return hess_dominant_c(n, H, ldh, lam_out, nullptr, u, nullptr);
}

}  // namespace imps

extern "C" {
// Host numerics of the Krylov drivers, exported for the CPU test suite.
int imps_host_sym_eig(int n, double* A, double* w, double* V) {
  imps::sym_eig(n, A, w, V);
  return 0;
}
int imps_host_hess_dominant(int n, const double* H, int ldh, double* lam, double* u) {
  return imps::hess_dominant(n, H, ldh, lam, u);
}
int imps_host_hess_eigvals(int n, const double* H, int ldh, double* wr, double* wi) {
  return imps::hess_eigvals(n, H, ldh, wr, wi);
}
// complex Hessenberg (re, im interleaved, column-major): dominant eigenvalue and unit eigenvector
int imps_host_hess_dominant_z(int n, const double* H, int ldh, double* lam, double* x, double* last) {
  std::complex<double> l;
  const int rc = imps::hess_dominant_z(n, reinterpret_cast<const std::complex<double>*>(H), ldh, &l,
                                       reinterpret_cast<std::complex<double>*>(x), last);
  lam[0] = l.real();
  lam[1] = l.imag();
  return rc;
}
}
