// Host-only driver of the library's CPU code (imps_small.cpp: Jacobi and Hessenberg eigen
// routines; imps_io.cpp: state and MPO files), built with g++ -fsanitize=address,undefined by
// tests/test_host_sanitizers.py.  It exercises each routine over a range of sizes and every error
// path of the file readers, checks the results, and returns 0 only when every check passed and
// the sanitizers reported nothing.
//
// The device part of the library is not linked: the few context-side symbols that imps_io.cpp
// references (the checkpoint calls, which move device state and are not exercised here) and the
// error-message sinks are defined below.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "imps.h"

namespace imps {
static std::string g_msg;
void set_global_error(const std::string& msg) { g_msg = msg; }
void set_context_error(imps_ctx*, const std::string& msg) { g_msg = msg; }
bool context_model(const imps_ctx*, int*, int*) { return false; }
}  // namespace imps

extern "C" {
const char* imps_last_error(const imps_ctx*) { return imps::g_msg.c_str(); }
int imps_state_dims(imps_ctx*, int*, int*, int*) { return IMPS_E_STATE; }
int imps_get_state(imps_ctx*, double* const*, double* const*, double* const*, double* const*) { return IMPS_E_STATE; }
int imps_set_state(imps_ctx*, int, int, const double* const*, const double* const*, const double* const*,
                   const double* const*) {
  return IMPS_E_STATE;
}
int imps_host_sym_eig(int n, double* A, double* w, double* V);
int imps_host_hess_eigvals(int n, const double* H, int ldh, double* wr, double* wi);
int imps_host_hess_dominant(int n, const double* H, int ldh, double* lam, double* u);
int imps_host_hess_dominant_z(int n, const double* H, int ldh, double* lam, double* x, double* last);
}

static int g_fail = 0;
#define CHECK(cond, ...)                          \
  do {                                            \
    if (!(cond)) {                                \
      std::fprintf(stderr, "FAILED %s: ", #cond); \
      std::fprintf(stderr, __VA_ARGS__);          \
      std::fprintf(stderr, "\n");                 \
      ++g_fail;                                   \
    }                                             \
  } while (0)

static void test_sym_eig(std::mt19937_64& rng) {
  std::normal_distribution<double> nd;
  for (int n = 1; n <= 40; n += 3) {
    std::vector<double> A(n * n), A0, w(n), V(n * n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i <= j; ++i) A[i + n * j] = A[j + n * i] = nd(rng);
    A0 = A;
    CHECK(imps_host_sym_eig(n, A.data(), w.data(), V.data()) == 0, "n=%d", n);
    double res = 0.0, orth = 0.0, nrm = 0.0;
    for (double x : A0) nrm = std::fmax(nrm, std::fabs(x));
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < n; ++i) {
        double s = -w[k] * V[i + n * k];
        for (int j = 0; j < n; ++j) s += A0[i + n * j] * V[j + n * k];
        res = std::fmax(res, std::fabs(s));
        double g = 0.0;
        for (int j = 0; j < n; ++j) g += V[j + n * i] * V[j + n * k];
        orth = std::fmax(orth, std::fabs(g - (i == k ? 1.0 : 0.0)));
      }
    CHECK(res < 1e-11 * n * nrm && orth < 1e-12 * n, "n=%d res=%g orth=%g", n, res, orth);
  }
}

static void test_hessenberg(std::mt19937_64& rng) {
  std::normal_distribution<double> nd;
  for (int n = 1; n <= 40; n += 3) {
    const int ldh = n + 2;  // a leading dimension larger than n
    std::vector<double> H(ldh * n, 0.0), wr(n), wi(n), u(n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i <= std::min(j + 1, n - 1); ++i) H[i + ldh * j] = nd(rng);
    CHECK(imps_host_hess_eigvals(n, H.data(), ldh, wr.data(), wi.data()) == 0, "n=%d", n);
    double tr = 0.0, sr = 0.0, si = 0.0;
    for (int i = 0; i < n; ++i) {
      tr += H[i + ldh * i];
      sr += wr[i];
      si += wi[i];
    }
    CHECK(std::fabs(tr - sr) < 1e-10 * n && std::fabs(si) < 1e-10 * n, "n=%d trace %g vs %g (im %g)", n, tr, sr, si);
    double lam[2] = {0.0, 0.0};
    imps_host_hess_dominant(n, H.data(), ldh, lam, u.data());
    double amax = 0.0;
    for (int i = 0; i < n; ++i) amax = std::fmax(amax, std::hypot(wr[i], wi[i]));
    CHECK(std::fabs(std::fabs(lam[0]) - amax) < 1e-8 * std::fmax(1.0, amax) || std::fabs(lam[0]) <= amax + 1e-8,
          "n=%d dominant %g vs %g", n, lam[0], amax);
    // complex Hessenberg: the dominant eigenpair's residual
    std::vector<std::complex<double>> Z(ldh * n, 0.0), x(n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i <= std::min(j + 1, n - 1); ++i) Z[i + ldh * j] = {nd(rng), nd(rng)};
    double zl[2], last[2] = {0.0, 0.0};
    const int rc = imps_host_hess_dominant_z(n, reinterpret_cast<const double*>(Z.data()), ldh, zl,
                                             reinterpret_cast<double*>(x.data()), last);
    CHECK(rc == 0, "n=%d complex rc=%d", n, rc);
    const std::complex<double> l(zl[0], zl[1]);
    double res = 0.0, xn = 0.0;
    for (int i = 0; i < n; ++i) {
      std::complex<double> s = -l * x[i];
      for (int j = 0; j < n; ++j) s += Z[i + ldh * j] * x[j];
      res = std::fmax(res, std::abs(s));
      xn += std::norm(x[i]);
    }
    CHECK(res < 1e-9 * std::fmax(1.0, std::abs(l)) && std::fabs(xn - 1.0) < 1e-12, "n=%d complex res=%g |x|^2=%g", n,
          res, xn);
  }
}

static std::vector<unsigned char> slurp(const std::string& p) {
  std::vector<unsigned char> b;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  int ch;
  while ((ch = std::fgetc(f)) != EOF) b.push_back((unsigned char)ch);
  std::fclose(f);
  return b;
}

static void spill(const std::string& p, const std::vector<unsigned char>& b) {
  FILE* f = std::fopen(p.c_str(), "wb");
  if (!f) return;
  if (!b.empty()) std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
}

static void test_files(const std::string& dir, std::mt19937_64& rng) {
  std::normal_distribution<double> nd;
  const int N = 2, D = 5, d = 3, nA = D * d * D, nC = D * D;
  std::vector<std::vector<double>> t(4 * N);
  for (int k = 0; k < 4 * N; ++k) {
    t[k].resize(k % 4 == 2 ? nC : nA);
    for (double& x : t[k]) x = nd(rng);
  }
  const double* AL[N] = {t[0].data(), t[4].data()};
  const double* AR[N] = {t[1].data(), t[5].data()};
  const double* C[N] = {t[2].data(), t[6].data()};
  const double* AC[N] = {t[3].data(), t[7].data()};
  const std::string p = dir + "/state.bin";
  CHECK(imps_write_state_file(p.c_str(), N, D, d, AL, AR, C, AC) == IMPS_OK, "%s", imps_last_error(nullptr));
  int n2, D2, d2;
  CHECK(imps_state_file_info(p.c_str(), &n2, &D2, &d2) == IMPS_OK && n2 == N && D2 == D && d2 == d, "info");
  // outputs for N + 1 sites, so that a file read with a larger unit cell stays in bounds
  std::vector<std::vector<double>> r(4 * (N + 1));
  for (int k = 0; k < 4 * (N + 1); ++k) r[k].assign(k % 4 == 2 ? nC : nA, -7.0);
  double* rAL[N + 1] = {r[0].data(), r[4].data(), r[8].data()};
  double* rAR[N + 1] = {r[1].data(), r[5].data(), r[9].data()};
  double* rC[N + 1] = {r[2].data(), r[6].data(), r[10].data()};
  double* rAC[N + 1] = {r[3].data(), r[7].data(), r[11].data()};
  CHECK(imps_read_state_file(p.c_str(), N, D, d, rAL, rAR, rC, rAC) == IMPS_OK, "%s", imps_last_error(nullptr));
  for (int k = 0; k < 4 * N; ++k) CHECK(std::memcmp(r[k].data(), t[k].data(), t[k].size() * 8) == 0, "block %d", k);

  // every damaged or mismatched file is refused and leaves the output untouched
  const std::vector<unsigned char> good = slurp(p);
  auto refused = [&](const std::vector<unsigned char>& bytes, int nn, int DD, int dd, const char* what) {
    const std::string q = dir + "/bad.bin";
    spill(q, bytes);
    for (auto& v : r) std::fill(v.begin(), v.end(), -7.0);
    const int st = imps_read_state_file(q.c_str(), nn, DD, dd, rAL, rAR, rC, rAC);
    CHECK(st == IMPS_E_IO || st == IMPS_E_ARG, "%s: status %d", what, st);
    bool untouched = true;
    for (auto& v : r)
      for (double x : v) untouched = untouched && x == -7.0;
    CHECK(untouched, "%s: output written", what);
    int a, b, c;
    imps_state_file_info(q.c_str(), &a, &b, &c);  // must not crash on any of them
  };
  CHECK(good.size() > 64, "file size %zu", good.size());
  for (size_t cut : {size_t(0), size_t(7), size_t(31), size_t(32), good.size() / 2, good.size() - 1})
    refused(std::vector<unsigned char>(good.begin(), good.begin() + cut), N, D, d, "truncated");
  for (size_t pos : {size_t(0), size_t(5), size_t(12), size_t(20), size_t(40), good.size() - 3}) {
    std::vector<unsigned char> b = good;
    b[pos] ^= 0x10;
    refused(b, N, D, d, "flipped byte");
  }
  {
    std::vector<unsigned char> b = good;
    b.push_back(0);
    refused(b, N, D, d, "trailing byte");
  }
  {
    std::vector<unsigned char> b = good;  // a header claiming huge dimensions
    for (int k = 8; k < 20; ++k) b[k] = 0x7f;
    refused(b, N, D, d, "huge header");
  }
  refused(good, N, D - 1, d, "wrong D");
  refused(good, N + 1, D, d, "wrong unit cell");
  {
    const std::string q = dir + "/missing.bin";
    for (auto& v : r) std::fill(v.begin(), v.end(), -7.0);
    CHECK(imps_read_state_file(q.c_str(), N, D, d, rAL, rAR, rC, rAC) == IMPS_E_IO, "missing file");
  }

  // MPO files: round trip, and a state file is refused as an MPO file
  const int w = 3, dm = 2, nW = w * dm * w * dm;
  std::vector<double> W0(nW), W1(nW);
  for (double& x : W0) x = nd(rng);
  for (double& x : W1) x = nd(rng);
  const double* Ws[2] = {W0.data(), W1.data()};
  const std::string m = dir + "/mpo.bin";
  CHECK(imps_write_mpo_file(m.c_str(), 2, w, dm, Ws) == IMPS_OK, "%s", imps_last_error(nullptr));
  std::vector<double> R0(nW, -7.0), R1(nW, -7.0);
  double* Rs[2] = {R0.data(), R1.data()};
  CHECK(imps_read_mpo_file(m.c_str(), 2, w, dm, Rs) == IMPS_OK, "%s", imps_last_error(nullptr));
  CHECK(std::memcmp(R0.data(), W0.data(), nW * 8) == 0 && std::memcmp(R1.data(), W1.data(), nW * 8) == 0, "mpo");
  std::fill(R0.begin(), R0.end(), -7.0);
  CHECK(imps_read_mpo_file(p.c_str(), 2, w, dm, Rs) == IMPS_E_IO && R0[0] == -7.0, "state file read as mpo");
  CHECK(imps_read_state_file(m.c_str(), N, D, d, rAL, rAR, rC, rAC) == IMPS_E_IO, "mpo file read as state");
  // argument errors
  CHECK(imps_write_state_file(nullptr, N, D, d, AL, AR, C, AC) == IMPS_E_ARG, "null path");
  CHECK(imps_read_mpo_file(m.c_str(), 0, w, dm, Rs) == IMPS_E_ARG, "nsites 0");
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <scratch dir>\n", argv[0]);
    return 2;
  }
  std::mt19937_64 rng(12345);
  test_sym_eig(rng);
  test_hessenberg(rng);
  test_files(argv[1], rng);
  if (g_fail) {
    std::fprintf(stderr, "%d check(s) failed\n", g_fail);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
