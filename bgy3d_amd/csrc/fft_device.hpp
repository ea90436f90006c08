// Device building blocks for the LDS-staged Stockham FFT (gfx950).
//
// A workgroup owns LINES lines of length N held in LDS.  Each stage reads
// R points per butterfly into registers, applies twiddles and an R-point
// DFT, synchronises, and writes them back in Stockham (autosort) order, so
// the result is in natural order without a bit-reversal pass.  N is a
// compile-time product of the radices 4, 2 and 3 (N = 2^a 3^b), which
// covers the grid sizes 32..384 the solver is built for (96 = 2^5 3 and
// 384 = 2^7 3 included).
//
// Sign convention: forward = exp(-2 pi i jk/N), unnormalised, like the
// FFTW plans the reference creates (bgy3d-fftw3.c:364-385).
#pragma once
#include <hip/hip_runtime.h>

namespace rism {

template <typename T> struct cpx { T re, im; };

template <typename T> __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T> __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T> __device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.re, -a.im}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> __device__ __forceinline__ cpx<T> rot90(cpx<T> a) {
  if constexpr (INV) return {-a.im, a.re};
  else return {a.im, -a.re};
}

// Radix list: 4s first, then one 2 if needed, then 3s.
__host__ __device__ constexpr int radix_at(int n, int s) {
  int k = 0;
  while (n % 4 == 0) { if (k == s) return 4; n /= 4; ++k; }
  while (n % 2 == 0) { if (k == s) return 2; n /= 2; ++k; }
  while (n % 3 == 0) { if (k == s) return 3; n /= 3; ++k; }
  return 0;
}
__host__ __device__ constexpr int num_stages(int n) {
  int k = 0;
  while (n % 4 == 0) { n /= 4; ++k; }
  while (n % 2 == 0) { n /= 2; ++k; }
  while (n % 3 == 0) { n /= 3; ++k; }
  return n == 1 ? k : -1;
}
__host__ __device__ constexpr int stage_ns(int n, int s) {  // product of earlier radices
  int p = 1;
  for (int i = 0; i < s; ++i) p *= radix_at(n, i);
  return p;
}

// LDS position of point p of a line, and the line stride (one point of
// padding, so consecutive lines start on different banks).  Padding inside
// the lines (one point per 16) was measured slower on gfx950: more
// SQ_LDS_BANK_CONFLICT cycles in the y and z passes, not fewer.
__host__ __device__ constexpr int lpos(int p) { return p; }
__host__ __device__ constexpr int lstride_for(int n) { return n + 1; }

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_small(cpx<T> (&v)[4]) {
  if constexpr (R == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cpx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cpx<T> a2 = cadd(v[1], v[3]), a3 = rot90<T, INV>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
  } else {  // R == 3
    const T h = T(0.86602540378443864676);  // sqrt(3)/2
    cpx<T> t1 = cadd(v[1], v[2]);
    cpx<T> t2 = {v[0].re - T(0.5) * t1.re, v[0].im - T(0.5) * t1.im};
    cpx<T> d = csub(v[1], v[2]);
    cpx<T> t3 = rot90<T, INV>(cpx<T>{d.re * h, d.im * h});
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
}

// One Stockham stage over the first nl of LINES lines of length N in LDS.
// Element (line l, point p) lives at buf[l * lstride + lpos(p)].
// tw[j] = exp(-2 pi i j / N), j < N (global memory, L1/L2 resident).
template <typename T, int N, int LINES, int NT, int STAGE, bool INV>
__device__ __forceinline__ void stockham_stage(cpx<T>* buf, int lstride, int nl,
                                               const cpx<T>* __restrict__ tw) {
  constexpr int R = radix_at(N, STAGE);
  constexpr int NS = stage_ns(N, STAGE);
  constexpr int NB = N / R;                 // butterflies per line
  constexpr int BF = LINES * NB;            // butterflies per workgroup
  constexpr int PER = (BF + NT - 1) / NT;   // per thread
  constexpr int TWS = N / (NS * R);         // twiddle stride
  cpx<T> v[PER][4];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    const int line = b / NB;
    if ((PER * NT == BF || b < BF) && line < nl) {
      const int j = b - line * NB;
      const int k = j % NS;
      cpx<T>* lb = buf + line * lstride;
#pragma unroll
      for (int q = 0; q < R; ++q) v[p][q] = lb[lpos(j + q * NB)];
      if constexpr (NS > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const cpx<T> w = tw[q * k * TWS];
          v[p][q] = INV ? cmulc(v[p][q], w) : cmul(v[p][q], w);
        }
      }
      dft_small<T, R, INV>(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    const int line = b / NB;
    if ((PER * NT == BF || b < BF) && line < nl) {
      const int j = b - line * NB;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
      cpx<T>* lb = buf + line * lstride;
#pragma unroll
      for (int q = 0; q < R; ++q) lb[lpos(base + q * NS)] = v[p][q];
    }
  }
  __syncthreads();
}

// FFT of the first nl (<= LINES) lines, each at buf + l * lstride.
template <typename T, int N, int LINES, int NT, bool INV, int S = 0>
__device__ __forceinline__ void fft_lines(cpx<T>* buf, int lstride, int nl,
                                          const cpx<T>* __restrict__ tw) {
  if constexpr (S < num_stages(N)) {
    stockham_stage<T, N, LINES, NT, S, INV>(buf, lstride, nl, tw);
    fft_lines<T, N, LINES, NT, INV, S + 1>(buf, lstride, nl, tw);
  }
}

// The same FFT on LINES lines stored interleaved: point p of line l at
// buf[p * LINES + l] (LINES a power of two), the layout in which rows of
// LINES consecutive kx columns arrive from global memory, so the copy in
// and out is a straight copy.  Butterflies are assigned line-fastest:
// consecutive lanes touch consecutive LDS words on every read and on every
// write after the first stage (whose writes are R-way strided).
template <typename T, int N, int LINES, int NT, int STAGE, bool INV>
__device__ __forceinline__ void stockham_stage_il(cpx<T>* buf, int nl,
                                                  const cpx<T>* __restrict__ tw) {
  static_assert((LINES & (LINES - 1)) == 0, "LINES must be a power of two");
  constexpr int R = radix_at(N, STAGE);
  constexpr int NS = stage_ns(N, STAGE);
  constexpr int NB = N / R;
  constexpr int BF = LINES * NB;
  constexpr int PER = (BF + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);
  cpx<T> v[PER][4];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    const int line = b & (LINES - 1);
    if ((PER * NT == BF || b < BF) && line < nl) {
      const int j = b / LINES;
      const int k = j % NS;
#pragma unroll
      for (int q = 0; q < R; ++q) v[p][q] = buf[(j + q * NB) * LINES + line];
      if constexpr (NS > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const cpx<T> w = tw[q * k * TWS];
          v[p][q] = INV ? cmulc(v[p][q], w) : cmul(v[p][q], w);
        }
      }
      dft_small<T, R, INV>(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    const int line = b & (LINES - 1);
    if ((PER * NT == BF || b < BF) && line < nl) {
      const int j = b / LINES;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[(base + q * NS) * LINES + line] = v[p][q];
    }
  }
  __syncthreads();
}

template <typename T, int N, int LINES, int NT, bool INV, int S = 0>
__device__ __forceinline__ void fft_lines_il(cpx<T>* buf, int nl, const cpx<T>* __restrict__ tw) {
  if constexpr (S < num_stages(N)) {
    stockham_stage_il<T, N, LINES, NT, S, INV>(buf, nl, tw);
    fft_lines_il<T, N, LINES, NT, INV, S + 1>(buf, nl, tw);
  }
}

// The FFT of the NCOL columns of a row-major plane in LDS: point p of
// column l at buf[p * LS + l], any row stride LS >= NCOL.  Butterflies are
// assigned column-fastest, so consecutive lanes touch consecutive LDS words
// (the layout of stockham_stage_il without its power-of-two line count).
template <typename T, int N, int LS, int NCOL, int NT, int STAGE, bool INV>
__device__ __forceinline__ void stockham_stage_cols(cpx<T>* buf, const cpx<T>* __restrict__ tw) {
  constexpr int R = radix_at(N, STAGE);
  constexpr int NS = stage_ns(N, STAGE);
  constexpr int NB = N / R;
  constexpr int BF = NCOL * NB;
  constexpr int PER = (BF + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);
  cpx<T> v[PER][4];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (PER * NT == BF || b < BF) {
      const int col = b % NCOL, j = b / NCOL;
      const int k = j % NS;
#pragma unroll
      for (int q = 0; q < R; ++q) v[p][q] = buf[(j + q * NB) * LS + col];
      if constexpr (NS > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          const cpx<T> w = tw[q * k * TWS];
          v[p][q] = INV ? cmulc(v[p][q], w) : cmul(v[p][q], w);
        }
      }
      dft_small<T, R, INV>(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int b = tid + p * NT;
    if (PER * NT == BF || b < BF) {
      const int col = b % NCOL, j = b / NCOL;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) buf[(base + q * NS) * LS + col] = v[p][q];
    }
  }
  __syncthreads();
}

template <typename T, int N, int LS, int NCOL, int NT, bool INV, int S = 0>
__device__ __forceinline__ void fft_cols(cpx<T>* buf, const cpx<T>* __restrict__ tw) {
  if constexpr (S < num_stages(N)) {
    stockham_stage_cols<T, N, LS, NCOL, NT, S, INV>(buf, tw);
    fft_cols<T, N, LS, NCOL, NT, INV, S + 1>(buf, tw);
  }
}

}  // namespace rism
