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

template <typename T, int R, bool INV, int NV>
__device__ __forceinline__ void dft_small(cpx<T> (&v)[NV]) {
  static_assert(R <= NV, "radix larger than the points held");
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
  } else if constexpr (R == 8) {
    // two 4-point DFTs of the even and odd points, then X_k = E_k + w^k O_k
    // and X_{k+4} = E_k - w^k O_k with w = exp(-+2 pi i / 8)
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<T, 4, INV>(e);
    dft_small<T, 4, INV>(o);
    const T h = T(0.70710678118654752440);  // 1/sqrt(2)
    // w o_1 = (1 -+ i) o_1 / sqrt 2, w^2 o_2 = -+i o_2, w^3 o_3 = (-1 -+ i) o_3 / sqrt 2
    const cpx<T> r1 = rot90<T, INV>(o[1]);
    const cpx<T> w1 = {h * (o[1].re + r1.re), h * (o[1].im + r1.im)};
    const cpx<T> w2 = rot90<T, INV>(o[2]);
    const cpx<T> r3 = rot90<T, INV>(o[3]);
    const cpx<T> w3 = {h * (r3.re - o[3].re), h * (r3.im - o[3].im)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], w1);
    v[5] = csub(e[1], w1);
    v[2] = cadd(e[2], w2);
    v[6] = csub(e[2], w2);
    v[3] = cadd(e[3], w3);
    v[7] = csub(e[3], w3);
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
template <typename T, int N, int LINES, int NT, int R, int NS, bool INV>
__device__ __forceinline__ void il_stage(cpx<T>* buf, int nl, const cpx<T>* __restrict__ tw) {
  static_assert((LINES & (LINES - 1)) == 0, "LINES must be a power of two");
  constexpr int NB = N / R;
  constexpr int BF = LINES * NB;
  constexpr int PER = (BF + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);
  cpx<T> v[PER][R];
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

template <typename T, int N, int LINES, int NT, int STAGE, bool INV>
__device__ __forceinline__ void stockham_stage_il(cpx<T>* buf, int nl,
                                                  const cpx<T>* __restrict__ tw) {
  il_stage<T, N, LINES, NT, radix_at(N, STAGE), stage_ns(N, STAGE), INV>(buf, nl, tw);
}

template <typename T, int N, int LINES, int NT, bool INV, int S = 0>
__device__ __forceinline__ void fft_lines_il(cpx<T>* buf, int nl, const cpx<T>* __restrict__ tw) {
  if constexpr (S < num_stages(N)) {
    stockham_stage_il<T, N, LINES, NT, S, INV>(buf, nl, tw);
    fft_lines_il<T, N, LINES, NT, INV, S + 1>(buf, nl, tw);
  }
}

// The interleaved FFT with the data held in registers on entry and exit.
// A thread holds R1 = LINES * N / NT points of one line l = tid % LINES,
// v[q] = point j0 + q * N / R1 with j0 = tid / LINES: the inputs of one
// radix-R1 butterfly of the first Stockham stage, which therefore runs on
// the registers and only writes LDS.  The later radices (8s while more than
// R1 points remain, the last one at most R1) run through LDS; the last
// stage reads LDS and keeps its outputs, which are again the points
// j0 + q * N / R1 of the thread's line (its R1 / RL butterflies cover
// exactly those), so v is overwritten in place with the transform.  Two LDS
// passes per stage after the first, against one copy in, two per stage and
// one copy out for fft_lines_il.
__host__ __device__ constexpr int rr_radix(int n, int r1, int s) {  // stage s >= 1
  int rem = n / r1;
  for (int i = 1;; ++i) {
    const int r = rem <= r1 ? rem : (rem % 8 == 0 && rem > 8 ? 8 : (rem % 4 == 0 && rem > 4 ? 4 : 2));
    if (i == s) return r;
    rem /= r;
    if (rem == 1) return 0;
  }
}
__host__ __device__ constexpr int rr_stages(int n, int r1) {
  int k = 1;
  while (rr_radix(n, r1, k) != 0) ++k;
  return k;
}
__host__ __device__ constexpr int rr_ns(int n, int r1, int s) {
  int p = r1;
  for (int i = 1; i < s; ++i) p *= rr_radix(n, r1, i);
  return p;
}
// Whether an N-point transform of LINES lines over NT threads keeps its
// points in registers: N a power of two, R1 = 4 or 8 points per thread
// exactly.  (R1 = 2 is supported by the index scheme but needs one stage
// more than fft_lines_il's radix-4 plan and more registers than it saves.)
__host__ __device__ constexpr bool rr_ok(int n, int lines, int nt) {
  if ((n & (n - 1)) != 0 || (lines * n) % nt != 0) return false;
  const int r1 = lines * n / nt;
  return (r1 == 4 || r1 == 8) && n > r1 && (nt % lines) == 0;
}

template <typename T, int N, int LINES, int NT, int R1, int S, bool INV>
__device__ __forceinline__ void rr_mid(cpx<T>* buf, int nl, const cpx<T>* __restrict__ tw) {
  if constexpr (S < rr_stages(N, R1) - 1) {
    il_stage<T, N, LINES, NT, rr_radix(N, R1, S), rr_ns(N, R1, S), INV>(buf, nl, tw);
    rr_mid<T, N, LINES, NT, R1, S + 1, INV>(buf, nl, tw);
  }
}

template <typename T, int N, int LINES, int NT, bool INV>
__device__ __forceinline__ void fft_lines_rr(cpx<T> (&v)[LINES * N / NT], cpx<T>* buf, int nl,
                                             const cpx<T>* __restrict__ tw) {
  constexpr int R1 = LINES * N / NT;
  static_assert(rr_ok(N, LINES, NT), "points per thread must be one radix-2, 4 or 8 butterfly");
  constexpr int S = rr_stages(N, R1);
  constexpr int RL = rr_radix(N, R1, S - 1);
  static_assert(RL >= 2 && R1 % RL == 0, "last radix must divide the first");
  const int tid = threadIdx.x;
  const int line = tid & (LINES - 1), j0 = tid / LINES;
  // stage 0 (NS = 1, no twiddles): outputs to points j0 R1 + q
  dft_small<T, R1, INV>(v);
  if (line < nl) {
#pragma unroll
    for (int q = 0; q < R1; ++q) buf[(j0 * R1 + q) * LINES + line] = v[q];
  }
  __syncthreads();
  rr_mid<T, N, LINES, NT, R1, 1, INV>(buf, nl, tw);
  // last stage (NS = N / RL): butterfly j = j0 + p N / R1 reads points
  // j + q N / RL, writes them back to the same points
  constexpr int NBL = N / RL;
#pragma unroll
  for (int p = 0; p < R1 / RL; ++p) {
    const int j = j0 + p * (N / R1);
    cpx<T> u[RL];
#pragma unroll
    for (int q = 0; q < RL; ++q) u[q] = buf[(j + q * NBL) * LINES + line];
#pragma unroll
    for (int q = 1; q < RL; ++q) {
      const cpx<T> w = tw[q * j];
      u[q] = INV ? cmulc(u[q], w) : cmul(u[q], w);
    }
    dft_small<T, RL, INV>(u);
#pragma unroll
    for (int q = 0; q < RL; ++q) v[p + q * (R1 / RL)] = u[q];
  }
  __syncthreads();  // the buffer is free for the next transform
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
