// 1D RISM / DRISM solvent susceptibility chi(k) = w(k) + rho h(k) (host C++).
//
// Replaces the reference's rism_solvent -> rism_vv (rism.f90:57-102,
// 310-583) for the solvent kernel of the 3D solver:
//   grid        mkgrid fft.f90:101-118, r_i = (i-1/2) dr, k_i = (i-1/2) dk,
//               dk = pi / rmax
//   transform   fourier_rows fft.f90:121-170 (DST-IV, FFTW RODFT11)
//   potentials  force_field rism.f90:2082-2162 (Lorentz-Berthelot, LJ +
//               erfc(alpha r)/r short range, 4 pi e^{-k^2/4a^2}/k^2 long)
//   bonds       omega_fourier rism.f90:1994-2032 (j0(k r_ab), bessel.f90)
//   DRISM       dipole_correction drism.f90:380-474
//   OZ          oz_vv_equation_c_t_MxM rism.f90:2476-2560
//   solve       r t(r) as unknowns (rism.f90:413-420), Newton-GMRES
// The DST-IV runs as a length-2n complex FFT of the pre-twiddled input;
// the Newton Jacobian is analytic.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "../../include/rism3d.h"

namespace {

using cd = std::complex<double>;
constexpr double PI = 3.14159265358979323846;
constexpr double EPS0INV = 331.84164;  // units.f90:50
constexpr double ALPHA = 1.2;          // units.f90:61

// ----------------------------------------------------------------- FFT
// Recursive mixed-radix (4, 2, 3, 5; other primes by direct DFT) complex
// FFT of length n over the twiddle table tw[j] = exp(sign 2 pi i j / N) of
// the top-level length N; unnormalised.
void fft_rec(const cd* in, cd* out, int n, int stride, const std::vector<cd>& tw, int N,
             std::vector<cd>& scratch) {
  if (n == 1) { out[0] = in[0]; return; }
  int r = 0;
  for (int c : {4, 2, 3, 5})
    if (n % c == 0) { r = c; break; }
  if (r == 0) {
    for (int p = 2; p * p <= n; ++p)
      if (n % p == 0) { r = p; break; }
    if (r == 0) r = n;
  }
  const int m = n / r;
  for (int q = 0; q < r; ++q) fft_rec(in + q * stride, out + q * m, m, stride * r, tw, N, scratch);
  const int sN = N / n, sR = N / r;
  cd tmp[64];
  std::vector<cd> big;
  cd* t = tmp;
  if (r > 64) { big.resize(r); t = big.data(); }
  for (int k = 0; k < m; ++k) {
    for (int q = 0; q < r; ++q) t[q] = out[q * m + k] * tw[(size_t)((long)q * k % n) * sN];
    for (int p = 0; p < r; ++p) {
      cd acc = 0;
      for (int q = 0; q < r; ++q) acc += t[q] * tw[(size_t)((long)p * q % r) * sR];
      scratch[p * m + k] = acc;
    }
  }
  for (int i = 0; i < n; ++i) out[i] = scratch[i];
}

struct Dst4 {
  int n;
  std::vector<cd> pre, post, a, A, scratch, tw;
  explicit Dst4(int n_) : n(n_), pre(n_), post(n_), a(2 * n_), A(2 * n_), scratch(2 * n_), tw(2 * n_) {
    for (int j = 0; j < n; ++j) pre[j] = std::polar(1.0, PI * j / (2.0 * n));
    for (int k = 0; k < n; ++k) post[k] = std::polar(1.0, PI * (2 * k + 1) / (4.0 * n));
    for (int j = 0; j < 2 * n; ++j) tw[j] = std::polar(1.0, 2 * PI * j / (2.0 * n));  // sign +1
  }
  // y_k = 2 sum_j x_j sin(pi (2j+1)(2k+1) / 4n)
  void operator()(const double* x, double* y) {
    for (int j = 0; j < n; ++j) a[j] = x[j] * pre[j];
    for (int j = n; j < 2 * n; ++j) a[j] = 0;
    fft_rec(a.data(), A.data(), 2 * n, 1, tw, 2 * n, scratch);
    for (int k = 0; k < n; ++k) y[k] = 2 * std::imag(post[k] * A[k]);
  }
};

// ------------------------------------------------------------ closures
// closures.f90:107-153 (PY has no differential in the reference).
double exps1(int n, double x) {
  double c = 1, y = 0;
  for (int k = n; k >= 1; --k) { y = x * y + c; c *= k; }
  return (x / c) * y;
}
double closure(int method, double bv, double t) {
  const double x = -bv + t;
  switch (method) {
    case 0: return std::expm1(x) - t;
    case 1: return x <= 0 ? std::expm1(x) - t : -bv;
    case 2: return std::exp(-bv) * (1 + t) - 1 - t;
    default: return (x < 0 ? std::expm1(x) : exps1(method - 3, x)) - t;
  }
}
double closure1(int method, double bv, double t) {  // dc/dt
  const double x = -bv + t;
  switch (method) {
    case 0: return std::expm1(x);
    case 1: return x <= 0 ? std::expm1(x) : 0.0;
    default: {
      const int n = method - 3 - 1;
      return x < 0 ? std::expm1(x) : exps1(n, x);
    }
  }
}

double sj0(double x) {
  if (std::fabs(x) < 0.5) {
    const double y = x * x;
    return 1 + y * (-1 / 6.0 + y * (1 / 120.0 + y * (-1 / 5040.0 + y * (1 / 362880.0 +
           y * (-1 / 39916800.0 + y * (1 / 6227020800.0))))));
  }
  return std::sin(x) / x;
}
double sj1(double x) {
  if (std::fabs(x) < 0.25) {
    const double y = x * x;
    return x / 3 * (1 + y * (-1 / 10.0 + y * (1 / 280.0 + y * (-1 / 15120.0 + y * (1 / 1330560.0 +
           y * (-1 / 172972800.0))))));
  }
  return (std::sin(x) / x - std::cos(x)) / x;
}

// Symmetric eigen-decomposition by cyclic Jacobi; columns of V are the
// eigenvectors (order and sign are immaterial below).
void jacobi_eig(int n, std::vector<double> A, std::vector<double>& V) {
  V.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (std::fabs(A[p * n + q]) < 1e-300) continue;
        const double th = (A[q * n + q] - A[p * n + p]) / (2 * A[p * n + q]);
        const double t = (th >= 0 ? 1 : -1) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
}

// Solve A X = B in place (A n x n, B n x nb), partial pivoting.
bool solve_small(int n, double* A, double* B, int nb) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
    if (A[p * n + c] == 0) return false;
    if (p != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[p * n + k]);
      for (int k = 0; k < nb; ++k) std::swap(B[c * nb + k], B[p * nb + k]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      for (int k = 0; k < nb; ++k) B[r * nb + k] -= f * B[c * nb + k];
    }
  }
  for (int r = n - 1; r >= 0; --r)
    for (int k = 0; k < nb; ++k) {
      double s = B[r * nb + k];
      for (int q = r + 1; q < n; ++q) s -= A[r * n + q] * B[q * nb + k];
      B[r * nb + k] = s / A[r * n + r];
    }
  return true;
}

struct Site3 {
  double x[3], sigma, eps, q;
};

// local coordinates of dipole_axes/local_coords (drism.f90:148-376)
std::vector<double> local_coords(const std::vector<Site3>& s) {
  const int m = (int)s.size();
  std::vector<double> w(m);
  double ws = 0;
  for (int i = 0; i < m; ++i) { w[i] = std::pow(s[i].sigma, 3); ws += w[i]; }
  for (auto& v : w) v /= ws;
  double c[3] = {0, 0, 0}, mu[3] = {0, 0, 0};
  for (int i = 0; i < m; ++i)
    for (int d = 0; d < 3; ++d) { c[d] += w[i] * s[i].x[d]; mu[d] += s[i].q * s[i].x[d]; }
  const double mn = std::sqrt(mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2]);
  for (double& v : mu) v /= mn;
  std::vector<double> T(9, 0.0), V;
  for (int i = 0; i < m; ++i) {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = s[i].x[k] - c[k];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) T[a * 3 + b] += w[i] * d[a] * d[b];
  }
  jacobi_eig(3, T, V);
  double v[3][3], e[3];
  for (int j = 0; j < 3; ++j) {
    double dotm = 0;
    for (int a = 0; a < 3; ++a) dotm += mu[a] * V[a * 3 + j];
    e[j] = 0;
    for (int a = 0; a < 3; ++a) { v[j][a] = V[a * 3 + j] - mu[a] * dotm; e[j] += v[j][a] * v[j][a]; }
  }
  const int imin = (int)(std::min_element(e, e + 3) - e);
  double u[3][3];
  int kk = 0;
  for (int j = 0; j < 3; ++j)
    if (j != imin) { for (int a = 0; a < 3; ++a) u[kk][a] = v[j][a]; ++kk; }
  for (int a = 0; a < 3; ++a) u[2][a] = mu[a];
  double d01 = 0, d00 = 0;
  for (int a = 0; a < 3; ++a) { d01 += u[0][a] * u[1][a]; d00 += u[0][a] * u[0][a]; }
  for (int a = 0; a < 3; ++a) u[1][a] -= u[0][a] * d01 / d00;
  for (int j = 0; j < 3; ++j) {
    double nn = 0;
    for (int a = 0; a < 3; ++a) nn += u[j][a] * u[j][a];
    nn = std::sqrt(nn);
    for (int a = 0; a < 3; ++a) u[j][a] /= nn;
  }
  std::vector<double> T2(4, 0.0), V2;
  for (int i = 0; i < m; ++i) {
    double dxy[2] = {0, 0};
    for (int j = 0; j < 2; ++j)
      for (int a = 0; a < 3; ++a) dxy[j] += (s[i].x[a] - c[a]) * u[j][a];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) T2[a * 2 + b] += w[i] * dxy[a] * dxy[b];
  }
  jacobi_eig(2, T2, V2);
  double un[2][3];
  for (int j = 0; j < 2; ++j)
    for (int a = 0; a < 3; ++a) un[j][a] = u[0][a] * V2[0 * 2 + j] + u[1][a] * V2[1 * 2 + j];
  for (int j = 0; j < 2; ++j)
    for (int a = 0; a < 3; ++a) u[j][a] = un[j][a];
  std::vector<double> x(3 * m);
  for (int i = 0; i < m; ++i)
    for (int d = 0; d < 3; ++d) {
      double sacc = 0;
      for (int a = 0; a < 3; ++a) sacc += u[d][a] * (s[i].x[a] - c[a]);
      x[d * m + i] = sacc;
    }
  return x;  // [3][m]
}

struct Rism1D {
  int m, n, method;
  double rmax, beta, rho, eps, dr, dk;
  std::vector<double> rs;  // number density of each site (rho for one molecule)
  std::vector<double> r, k;
  std::vector<double> vr, vk, wk, xk, W;  // [m][m][n]
  Dst4 dst;
  Rism1D(int m_, int n_, int method_, double rmax_, double beta_, double rho_, double eps_)
      : m(m_), n(n_), method(method_), rmax(rmax_), beta(beta_), rho(rho_), eps(eps_), rs(m_, rho_),
        dst(n_) {}

  size_t at(int i, int j, int p) const { return ((size_t)i * m + j) * n + p; }

  // fourier_rows: g_p = DST4[f_i 2n(2i-1)]_p / (2p-1), scaled by s
  void ft(const double* f, double* g, double s, std::vector<double>& tmp) {
    tmp.resize(n);
    for (int i = 0; i < n; ++i) tmp[i] = f[i] * (2.0 * n * (2 * i + 1));
    dst(tmp.data(), g);
    for (int p = 0; p < n; ++p) g[p] = g[p] / (2 * p + 1) * s;
  }

  // t(k) = OZ(c(k)) per k: H = (1 - W C D)^-1 W C W, D = diag(rs)
  void oz(const std::vector<double>& c, std::vector<double>& t) const {
    std::vector<double> WC(m * m), A(m * m), H(m * m);
    for (int p = 0; p < n; ++p) {
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          double s = 0;
          for (int q = 0; q < m; ++q) s += W[at(i, q, p)] * c[at(q, j, p)];
          WC[i * m + j] = s;
        }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          A[i * m + j] = (i == j ? 1.0 : 0.0) - WC[i * m + j] * rs[j];  // W C D
          double s = 0;
          for (int q = 0; q < m; ++q) s += WC[i * m + q] * W[at(q, j, p)];
          H[i * m + j] = s;
        }
      std::vector<double> AA = A;
      solve_small(m, AA.data(), H.data(), m);
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) t[at(i, j, p)] = H[i * m + j] - c[at(i, j, p)];
    }
  }
  // dT = A^-1 W dC (W + D H) - dC
  void oz1(const std::vector<double>& c, const std::vector<double>& dc, std::vector<double>& dt) const {
    std::vector<double> WC(m * m), A(m * m), H(m * m), X(m * m), Y(m * m);
    for (int p = 0; p < n; ++p) {
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          double s = 0;
          for (int q = 0; q < m; ++q) s += W[at(i, q, p)] * c[at(q, j, p)];
          WC[i * m + j] = s;
        }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          A[i * m + j] = (i == j ? 1.0 : 0.0) - WC[i * m + j] * rs[j];  // W C D
          double s = 0;
          for (int q = 0; q < m; ++q) s += WC[i * m + q] * W[at(q, j, p)];
          H[i * m + j] = s;
        }
      std::vector<double> AA = A;
      solve_small(m, AA.data(), H.data(), m);  // H = A^-1 W C W
      // X = W dC, Y = X (W + rho H)
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          double s = 0;
          for (int q = 0; q < m; ++q) s += W[at(i, q, p)] * dc[at(q, j, p)];
          X[i * m + j] = s;
        }
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          double s = 0;
          for (int q = 0; q < m; ++q) s += X[i * m + q] * (W[at(q, j, p)] + rs[q] * H[q * m + j]);
          Y[i * m + j] = s;
        }
      AA = A;
      solve_small(m, AA.data(), Y.data(), m);
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) dt[at(i, j, p)] = Y[i * m + j] - dc[at(i, j, p)];
    }
  }

  // F(rt) = r [T(t) - t], t = rt / r  (iterate_rt, rism.f90:539-546)
  void F(const std::vector<double>& rt, std::vector<double>& out, std::vector<double>* cks = nullptr) {
    const size_t N = (size_t)m * m * n;
    std::vector<double> c(N), ck(N), tt(N), tmp;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        for (int p = 0; p < n; ++p) {
          const size_t q = at(i, j, p);
          c[q] = closure(method, beta * vr[q], rt[q] / r[p]);
        }
    for (int i = 0; i < m * m; ++i) ft(&c[(size_t)i * n], &ck[(size_t)i * n], dr * dr * dr, tmp);
    for (size_t q = 0; q < N; ++q) ck[q] -= beta * vk[q];
    oz(ck, tt);
    for (size_t q = 0; q < N; ++q) tt[q] += xk[q] - beta * vk[q];
    const double bw = dk * dk * dk / std::pow(2 * PI, 3);
    out.assign(N, 0.0);
    for (int i = 0; i < m * m; ++i) ft(&tt[(size_t)i * n], &out[(size_t)i * n], bw, tmp);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        for (int p = 0; p < n; ++p) {
          const size_t q = at(i, j, p);
          out[q] = r[p] * (out[q] - rt[q] / r[p]);
        }
    if (cks) *cks = ck;
  }
  void J(const std::vector<double>& rt, const std::vector<double>& ck, const std::vector<double>& drt,
         std::vector<double>& out) {
    const size_t N = (size_t)m * m * n;
    std::vector<double> dc(N), dck(N), dtt(N), tmp;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        for (int p = 0; p < n; ++p) {
          const size_t q = at(i, j, p);
          dc[q] = closure1(method, beta * vr[q], rt[q] / r[p]) * drt[q] / r[p];
        }
    for (int i = 0; i < m * m; ++i) ft(&dc[(size_t)i * n], &dck[(size_t)i * n], dr * dr * dr, tmp);
    oz1(ck, dck, dtt);
    const double bw = dk * dk * dk / std::pow(2 * PI, 3);
    out.assign(N, 0.0);
    for (int i = 0; i < m * m; ++i) ft(&dtt[(size_t)i * n], &out[(size_t)i * n], bw, tmp);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        for (int p = 0; p < n; ++p) {
          const size_t q = at(i, j, p);
          out[q] = r[p] * (out[q] - drt[q] / r[p]);
        }
  }
};

double dotv(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
double maxabs(const std::vector<double>& a) {
  double s = 0;
  for (double v : a) s = std::max(s, std::fabs(v));
  return s;
}

// GMRES(restart) for J x = b from x = 0, relative tolerance rtol.
template <class Op>
void gmres(Op A, const std::vector<double>& b, std::vector<double>& x, double rtol, int restart, int maxit) {
  const size_t N = b.size();
  x.assign(N, 0.0);
  const double bn = std::sqrt(dotv(b, b));
  if (bn == 0) return;
  std::vector<double> r = b, w;
  int it = 0;
  while (it < maxit) {
    const double beta0 = std::sqrt(dotv(r, r));
    if (beta0 <= rtol * bn) return;
    std::vector<std::vector<double>> V(restart + 1, std::vector<double>(N));
    std::vector<double> H((restart + 1) * restart, 0.0), cs(restart), sn(restart), g(restart + 1, 0.0);
    for (size_t i = 0; i < N; ++i) V[0][i] = r[i] / beta0;
    g[0] = beta0;
    int kdim = 0;
    for (int j = 0; j < restart && it < maxit; ++j, ++it) {
      A(V[j], w);
      for (int i = 0; i <= j; ++i) {  // modified Gram-Schmidt
        const double h = dotv(w, V[i]);
        H[i * restart + j] = h;
        for (size_t q = 0; q < N; ++q) w[q] -= h * V[i][q];
      }
      const double hn = std::sqrt(dotv(w, w));
      H[(j + 1) * restart + j] = hn;
      if (hn > 0)
        for (size_t q = 0; q < N; ++q) V[j + 1][q] = w[q] / hn;
      for (int i = 0; i < j; ++i) {
        const double a = H[i * restart + j], bb = H[(i + 1) * restart + j];
        H[i * restart + j] = cs[i] * a + sn[i] * bb;
        H[(i + 1) * restart + j] = -sn[i] * a + cs[i] * bb;
      }
      const double a = H[j * restart + j], bb = H[(j + 1) * restart + j];
      const double den = std::sqrt(a * a + bb * bb);
      cs[j] = den > 0 ? a / den : 1;
      sn[j] = den > 0 ? bb / den : 0;
      H[j * restart + j] = den;
      H[(j + 1) * restart + j] = 0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      kdim = j + 1;
      if (std::fabs(g[j + 1]) <= rtol * bn || hn == 0) break;
    }
    std::vector<double> y(kdim);
    for (int i = kdim - 1; i >= 0; --i) {
      double s = g[i];
      for (int q = i + 1; q < kdim; ++q) s -= H[i * restart + q] * y[q];
      y[i] = s / H[i * restart + i];
    }
    for (int i = 0; i < kdim; ++i)
      for (size_t q = 0; q < N; ++q) x[q] += y[i] * V[i][q];
    A(x, w);
    for (size_t q = 0; q < N; ++q) r[q] = b[q] - w[q];
    if (std::fabs(g[kdim]) <= rtol * bn) return;
  }
}

}  // namespace

namespace {
// One molecule at pd->rho with DRISM (molecule == nullptr), or a mixture
// (XRISM): molecule[i] the molecule of site i, rho_sites[i] its density.
int solvent_chi(const rism3d_problem* pd, int m, const rism3d_site* solvent, double dielectric,
                const int* molecule, const double* rho_sites, double* chi_out) {
  if (!pd || !solvent || !chi_out || m < 1 || m > 8 || pd->nrad < 4 || !(pd->rmax > 0) ||
      !(pd->beta > 0) || !(pd->rho > 0) || pd->closure < 0 || pd->closure > 10)
    return RISM3D_BAD_ARG;
  if (molecule) {
    if (!rho_sites || dielectric != 0) return RISM3D_BAD_ARG;  // mixtures: XRISM only
    for (int i = 0; i < m; ++i) {
      if (!(rho_sites[i] > 0)) return RISM3D_BAD_ARG;
      for (int j = 0; j < m; ++j)  // sites of one molecule share its density
        if (molecule[i] == molecule[j] && rho_sites[i] != rho_sites[j]) return RISM3D_BAD_ARG;
    }
  }
  const int n = pd->nrad;
  std::vector<Site3> s(m);
  for (int i = 0; i < m; ++i) {
    for (int d = 0; d < 3; ++d) s[i].x[d] = solvent[i].x[d];
    s[i].sigma = solvent[i].sigma;
    s[i].eps = solvent[i].epsilon;
    s[i].q = solvent[i].charge;
  }
  Rism1D R(m, n, pd->closure, pd->rmax, pd->beta, pd->rho, dielectric);
  if (molecule)
    for (int i = 0; i < m; ++i) R.rs[i] = rho_sites[i];
  R.dr = pd->rmax / n;
  R.dk = PI / pd->rmax;
  R.r.resize(n);
  R.k.resize(n);
  for (int i = 0; i < n; ++i) {
    R.r[i] = (2 * (i + 1) - 1) * R.dr / 2;
    R.k[i] = (2 * (i + 1) - 1) * R.dk / 2;
  }
  const size_t N = (size_t)m * m * n;
  R.vr.assign(N, 0);
  R.vk.assign(N, 0);
  R.wk.assign(N, 0);
  R.xk.assign(N, 0);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const double qab = s[i].q * s[j].q;
      const double sig = (s[i].sigma + s[j].sigma) / 2;
      const double eps = std::sqrt(s[i].eps * s[j].eps);
      double rab = 0;
      for (int d = 0; d < 3; ++d) rab += (s[i].x[d] - s[j].x[d]) * (s[i].x[d] - s[j].x[d]);
      rab = std::sqrt(rab);
      for (int p = 0; p < n; ++p) {
        const double rr = R.r[p], kk = R.k[p];
        double v = EPS0INV * qab * std::erfc(ALPHA * rr) / rr;
        if (sig != 0) {
          const double sr6 = 1 / std::pow(rr / sig, 6);
          v += eps * 4 * sr6 * (sr6 - 1);
        }
        R.vr[R.at(i, j, p)] = v;
        R.vk[R.at(i, j, p)] = EPS0INV * qab * 4 * PI * std::exp(-kk * kk / (4 * ALPHA * ALPHA)) / (kk * kk);
        R.wk[R.at(i, j, p)] = (molecule && molecule[i] != molecule[j]) ? 0.0 : sj0(kk * rab);
      }
    }
  if (dielectric != 0) {
    double mu[3] = {0, 0, 0};
    for (int i = 0; i < m; ++i)
      for (int d = 0; d < 3; ++d) mu[d] += s[i].q * s[i].x[d];
    const double mu2 = mu[0] * mu[0] + mu[1] * mu[1] + mu[2] * mu[2];
    if (!(mu2 > 0)) return RISM3D_BAD_ARG;  // DRISM needs a dipolar solvent
    const double y = 4 * PI * pd->beta * EPS0INV * pd->rho * mu2 / 9;
    const std::vector<double> x = local_coords(s);
    for (int p = 0; p < n; ++p) {
      const double kk = R.k[p];
      const double h = (dielectric - (1 + 3 * y)) / (pd->rho * y) * std::exp(-(kk * 0.5) * (kk * 0.5) / 4);
      std::vector<double> f(m);
      for (int i = 0; i < m; ++i) f[i] = sj0(kk * x[0 * m + i]) * sj0(kk * x[1 * m + i]) * sj1(kk * x[2 * m + i]);
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) R.xk[R.at(i, j, p)] = f[i] * h * f[j];
    }
  }
  R.W.resize(N);
  for (size_t q = 0; q < N; ++q) R.W[q] = R.wk[q] + pd->rho * R.xk[q];

  // Newton-GMRES with backtracking on 0.5 |F|^2 from r t = 0
  std::vector<double> rt(N, 0.0), f, fn, ck, dx, xn;
  R.F(rt, f, &ck);
  const double tol = 1e-12;
  int it = 0;
  for (; it < 100 && maxabs(f) > tol; ++it) {
    if (!std::isfinite(maxabs(f))) return RISM3D_NAN;
    std::vector<double> negf(N);
    for (size_t q = 0; q < N; ++q) negf[q] = -f[q];
    if (pd->closure == 2) {  // PY: finite-difference J.v (the reference has no closure1 for PY)
      gmres([&](const std::vector<double>& v, std::vector<double>& out) {
              const double nv = std::sqrt(dotv(v, v));
              if (nv == 0) { out.assign(N, 0.0); return; }
              const double e = 1e-7 * (1 + std::sqrt(dotv(rt, rt))) / nv;
              std::vector<double> z(N), fz;
              for (size_t q = 0; q < N; ++q) z[q] = rt[q] + e * v[q];
              R.F(z, fz);
              out.resize(N);
              for (size_t q = 0; q < N; ++q) out[q] = (fz[q] - f[q]) / e;
            }, negf, dx, 1e-4, 30, 300);
    } else {
      gmres([&](const std::vector<double>& v, std::vector<double>& out) { R.J(rt, ck, v, out); },
            negf, dx, 1e-4, 30, 300);
    }
    const double f0 = 0.5 * dotv(f, f);
    double lam = 1;
    std::vector<double> ckn;
    for (;;) {
      xn = rt;
      for (size_t q = 0; q < N; ++q) xn[q] += lam * dx[q];
      R.F(xn, fn, &ckn);
      const double f1 = 0.5 * dotv(fn, fn);
      if (std::isfinite(f1) && f1 <= (1 - 1e-4 * lam) * f0) break;
      lam *= 0.5;
      if (lam < 1e-6) break;
    }
    rt.swap(xn);
    f.swap(fn);
    ck.swap(ckn);
  }
  if (!(maxabs(f) <= tol)) return std::isfinite(maxabs(f)) ? RISM3D_NONCONVERGED : RISM3D_NAN;

  // chi = w + rho FT(c + t) dr^3 (rism.f90:441-449)
  std::vector<double> ct(N), h(N), tmp;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      for (int p = 0; p < n; ++p) {
        const size_t q = R.at(i, j, p);
        const double t = rt[q] / R.r[p];
        ct[q] = closure(pd->closure, pd->beta * R.vr[q], t) + t;
      }
  for (int i = 0; i < m * m; ++i) R.ft(&ct[(size_t)i * n], &h[(size_t)i * n], R.dr * R.dr * R.dr, tmp);
  // chi_ij = w_ij + rho_i h_ij
  for (int i = 0; i < m; ++i)
    for (size_t q = (size_t)i * m * n; q < (size_t)(i + 1) * m * n; ++q)
      chi_out[q] = R.wk[q] + R.rs[i] * h[q];
  return RISM3D_OK;
}
}  // namespace

extern "C" int rism3d_solvent_chi_internal(const rism3d_problem* pd, int m,
                                           const rism3d_site* solvent, double dielectric,
                                           double* chi_out) {
  return solvent_chi(pd, m, solvent, dielectric, nullptr, nullptr, chi_out);
}

extern "C" int rism3d_solvent_chi_mixture(const rism3d_problem* pd, int m,
                                          const rism3d_site* solvent, const int* molecule,
                                          const double* rho_sites, double* chi_out) {
  if (!molecule) return RISM3D_BAD_ARG;
  return solvent_chi(pd, m, solvent, 0.0, molecule, rho_sites, chi_out);
}
