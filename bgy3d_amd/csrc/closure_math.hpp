// Pointwise closure functions shared by the kernels (device) and the
// host-side unit test of their rounding (tests/test_closure_math.py, which
// compiles this header with g++ and RISM_HD defined as inline).
#pragma once
#include <cmath>

#ifndef RISM_HD
#define RISM_HD __host__ __device__ __forceinline__
#endif

namespace rism {

template <typename T> RISM_HD T d_expm1(T x);
template <> RISM_HD double d_expm1<double>(double x) { return std::expm1(x); }
template <> RISM_HD float d_expm1<float>(float x) { return std::expm1(x); }
template <typename T> RISM_HD T d_exp(T x);
template <> RISM_HD double d_exp<double>(double x) { return std::exp(x); }
template <> RISM_HD float d_exp<float>(float x) { return std::exp(x); }

// Truncated exponential series sum_{k=1..n} x^k/k! by Horner
// (closures.f90:191-236).
template <typename T> RISM_HD T exps1(int n, T x) {
  T c = T(1), y = T(0);
  for (int k = n; k >= 1; --k) { y = x * y + c; c = c * T(k); }
  return (x / c) * y;
}

// c = closure(beta v, t) (closures.f90:107-128).  bv = beta * v.
template <typename T> RISM_HD T closure_c(int method, int order, T bv, T t) {
  const T x = -bv + t;
  switch (method) {
    case 0:  // HNC
      return d_expm1(x) - t;
    case 1:  // KH
      return (x <= T(0)) ? d_expm1(x) - t : -bv;
    case 2:  // PY
      return d_exp(-bv) * (T(1) + t) - T(1) - t;
    default:  // PSE-n
      return ((x < T(0)) ? d_expm1(x) : exps1(order, x)) - t;
  }
}

// c(t_b + d) - c(t_b) for the closures of the form c = f(x) - t,
// x = -beta v + t (all but PY), from x_b and d, in a form whose rounding
// error is relative to the difference rather than to f: on the exponential
// branch f(x_b + d) - f(x_b) = e^{x_b} expm1(d); on the PSE-n polynomial
// branch sum_{j=1..n} d^j/j! S_{n-j}(x_b) with S_k the k-th partial sum of
// e^x; KH's linear branch d.  Where the step crosses the branch point both
// values are small (|x_b| < |d|) and are subtracted directly.
template <typename T> RISM_HD T closure_f(int method, int order, T x) {
  switch (method) {
    case 0: return d_expm1(x);
    case 1: return (x <= T(0)) ? d_expm1(x) : x;
    default: return (x < T(0)) ? d_expm1(x) : exps1(order, x);
  }
}
// e^{x_b + d} - e^{x_b} as e^{max} times an expm1 in (-1, 0]: with the
// larger exponent outside, the product neither overflows nor forms 0 * inf
// (deep in a repulsive core x_b ~ -1e9 while d may be large).
// Written without branches around the exponentials: with d of either sign
// in one wavefront both sides of a branch would run.  d > 0: -e^{x_b + d}
// expm1(-d); else e^{x_b} expm1(d) -- the same operations either way.
template <typename T> RISM_HD T exp_diff(T xb, T d) {
  const T y = xb + d;
  const T e = d_exp(d > T(0) ? y : xb);
  const T m = d_expm1(d > T(0) ? -d : d);
  return (d > T(0)) ? -e * m : e * m;
}
template <typename T>
RISM_HD T closure_dc(int method, int order, T xb, T d) {
  const T y = xb + d;
  T df;
  if (method == 0) {
    df = exp_diff(xb, d);
  } else if (method == 1) {
    // both branches of f on the exponential side: exp_diff; both linear: d;
    // across the branch point the linear side is exact and the other one
    // expm1 of the smaller argument -- one exp and two expm1 per point
    const T ed = exp_diff(xb, d);
    const T em = d_expm1(xb < y ? xb : y);
    if (xb <= T(0) && y <= T(0)) df = ed;
    else if (xb > T(0) && y > T(0)) df = d;
    else df = (xb <= T(0)) ? y - em : em - xb;
  } else {
    if (xb < T(0) && y < T(0)) {
      df = exp_diff(xb, d);
    } else if (xb >= T(0) && y >= T(0)) {
      // S_k(x) = sum_{i<=k} x^i/i!; df = sum_j d^j/j! S_{order-j}(x_b)
      T S[8];
      T term = T(1);
      S[0] = T(1);
      for (int k = 1; k < order; ++k) { term = term * xb / T(k); S[k] = S[k - 1] + term; }
      T dj = T(1);
      df = T(0);
      for (int j = 1; j <= order; ++j) {
        dj = dj * d / T(j);
        df += dj * S[order - j];
      }
    } else {
      df = closure_f<T>(method, order, y) - closure_f<T>(method, order, xb);
    }
  }
  return df - d;
}

}  // namespace rism
