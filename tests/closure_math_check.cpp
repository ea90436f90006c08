// Host check of closure_dc (closure_math.hpp): the FP32 and FP64 closure
// differences against a long-double reference f(x_b + d) - f(x_b) - d.
// Prints, per closure, the largest error relative to the size of the step,
// |f(x_b + d) - f(x_b)| + |d|, for closure_dc<float>, closure_dc<double>
// and the plain FP32 difference of closure_f.
#include <cmath>
#include <cstdio>
#include <initializer_list>
#define RISM_HD inline
#include "../bgy3d_amd/csrc/closure_math.hpp"

namespace rism {
template <> inline long double d_expm1<long double>(long double x) { return expm1l(x); }
template <> inline long double d_exp<long double>(long double x) { return expl(x); }
}  // namespace rism

int main() {
  const int methods[][2] = {{0, 0}, {1, 1}, {5, 2}, {6, 3}, {9, 6}};
  const char* names[] = {"HNC", "KH", "PSE2", "PSE3", "PSE6"};
  for (int mi = 0; mi < 5; ++mi) {
    const int m = methods[mi][0], n = methods[mi][1];
    double e32 = 0, e64 = 0, enaive = 0;
    for (int i = 0; i <= 400; ++i) {
      const float xb = -30.0f + 0.1f * i + 0.0137f;  // [-30, 10]
      for (int j = -60; j <= 60; ++j) {
        if (j == 0) continue;
        const float d = (j > 0 ? 1.f : -1.f) * std::pow(10.f, -0.1f * std::abs(j));  // 1e-6..1
        const long double ref = rism::closure_f<long double>(m, n, (long double)xb + d) -
                                rism::closure_f<long double>(m, n, (long double)xb) - d;
        const long double den = std::fabs(ref + d) + std::fabs((long double)d);
        const double a32 = rism::closure_dc<float>(m, n, xb, d);
        const double a64 = rism::closure_dc<double>(m, n, (double)xb, (double)d);
        const float naive = rism::closure_f<float>(m, n, xb + d) - rism::closure_f<float>(m, n, xb) - d;
        e32 = std::fmax(e32, (double)(std::fabs(a32 - ref) / den));
        e64 = std::fmax(e64, (double)(std::fabs(a64 - ref) / den));
        enaive = std::fmax(enaive, (double)(std::fabs(naive - ref) / den));
      }
    }
    // deep in a repulsive core: x_b ~ -1e9, steps up to 1e3, must stay finite
    int nonfinite = 0;
    for (float xb : {-1.9e9f, -1e4f, -200.f, -90.f}) {
      for (float d : {-1e3f, -100.f, -1.f, 1.f, 100.f, 1e3f}) {
        if (xb + d > 0) continue;
        if (!std::isfinite(rism::closure_dc<float>(m, n, xb, d))) ++nonfinite;
        if (!std::isfinite(rism::closure_dc<double>(m, n, (double)xb, (double)d))) ++nonfinite;
      }
    }
    std::printf("%s %.3e %.3e %.3e %d\n", names[mi], e32, e64, enaive, nonfinite);
  }
  return 0;
}
