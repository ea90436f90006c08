// gfx950 kernels of the 3D-RISM residual F(t) = (chi - 1) * c(t) - t and
// its setup / post-processing.  One residual evaluation is five launches:
//
//   1. closure_r2c_x : c = closure(beta v, t) fused into the x-axis r2c
//   2. fft_y         : forward c2c along y
//   3. zstar         : forward c2c along z, per-k m x m chi contraction,
//                      tau (or Ng long-range) term, inverse c2c along z
//   4. fft_y         : inverse c2c along y
//   5. c2r_x_resid   : x-axis c2r fused with r = t' - t and its norms
//
// References: iterate_t1 hnc3d.c:1184-1285, star hnc3d.c:1115-1138,
// compute_c hnc3d.c:167-181, closures.f90:107-128 and 248-432,
// interp bgy3d-vec.c:23-46, KFREQ bgy3d.h:162, FFT layout
// bgy3d-fftw3.c:364-385.
#include "fft_device.hpp"
#include "kernels.hpp"
#include "closure_math.hpp"

namespace rism {

// closure_c, closure_dc and their helpers: closure_math.hpp

// Chemical-potential density form(x, h, c, cl) (closures.f90:574-612).
__device__ __forceinline__ double form_mu(int method, int order, double x, double h, double c,
                                          double cl) {
  const double base = -c - h * (c + cl) / 2;
  switch (method) {
    case 0: return base + h * h / 2;
    case 1: return (h < 0) ? base + h * h / 2 : base;
    case 2: return base;
    default: {
      if (h < 0) return base + h * h / 2;
      const int m = order + 1;
      double mx = 1, xm = 1;
      for (int k = 1; k <= m; ++k) { mx *= k; xm *= x; }
      return base + (h * h / 2 - xm / mx);
    }
  }
}

// ------------------------------------------------------------- row sizes
__host__ __device__ constexpr int rows_for(int M) {
  return M <= 16 ? 64 : M <= 32 ? 32 : M <= 64 ? 16 : M <= 128 ? 8 : 4;
}

// 16-byte vectors of the working type: the streaming loops of the x-axis
// kernels move 16 B per lane per array (global_load/store_dwordx4), four
// times fewer memory instructions than scalar accesses.  Every array is
// addressed at multiples of N0 >= 32 elements from a 256-byte aligned base.
template <typename T> struct alignas(16) Vec16 {
  static constexpr int W = 16 / sizeof(T);
  T v[W];
};
// Global-memory accessors.  The array pointers come out of descriptors in
// memory, so the compiler sees generic pointers and emits flat accesses,
// which also count against the LDS wait counter: every wait for an LDS
// access would then wait for the outstanding HBM loads too.  These casts to
// address space 1 (through native vector types; aggregate copies lose the
// address space) give global_load / global_store.
template <typename T> struct VecOf;
template <> struct VecOf<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef float v2 __attribute__((ext_vector_type(2)));
};
template <> struct VecOf<double> {
  typedef double v4 __attribute__((ext_vector_type(2)));  // 16 bytes
  typedef double v2 __attribute__((ext_vector_type(2)));
};
template <typename T> __device__ __forceinline__ Vec16<T> ld16(const T* p) {
  typedef typename VecOf<T>::v4 V;
  typedef __attribute__((address_space(1))) const V GV;
  const V x = *(GV*)p;
  Vec16<T> r;
#pragma unroll
  for (int k = 0; k < Vec16<T>::W; ++k) r.v[k] = x[k];
  return r;
}
template <typename T> __device__ __forceinline__ void st16(T* p, const Vec16<T>& r) {
  typedef typename VecOf<T>::v4 V;
  typedef __attribute__((address_space(1))) V GV;
  V x;
#pragma unroll
  for (int k = 0; k < Vec16<T>::W; ++k) x[k] = r.v[k];
  *(GV*)p = x;
}
template <typename T> __device__ __forceinline__ cpx<T> ldc(const cpx<T>* p) {
  typedef typename VecOf<T>::v2 V;
  typedef __attribute__((address_space(1))) const V GV;
  const V x = *(GV*)p;
  return cpx<T>{x[0], x[1]};
}
template <typename T> __device__ __forceinline__ void stc(cpx<T>* p, cpx<T> c) {
  typedef typename VecOf<T>::v2 V;
  typedef __attribute__((address_space(1))) V GV;
  V x;
  x[0] = c.re;
  x[1] = c.im;
  *(GV*)p = x;
}
// base[i] for a base the same in every lane and an element index below
// 2^32 bytes from it: the byte offset stays 32-bit, so the access takes the
// scalar-base form instead of a 64-bit address per lane.
template <typename X> __device__ __forceinline__ X* at32(X* base, uint32_t i) {
  return reinterpret_cast<X*>(reinterpret_cast<char*>(base) + i * (uint32_t)sizeof(X));
}
template <typename X> __device__ __forceinline__ const X* at32(const X* base, uint32_t i) {
  return reinterpret_cast<const X*>(reinterpret_cast<const char*>(base) + i * (uint32_t)sizeof(X));
}
template <typename X> __device__ __forceinline__ X gld(const X* p) {  // scalars
  typedef __attribute__((address_space(1))) const X GX;
  return *(GX*)p;
}
template <typename X> __device__ __forceinline__ void gst(X* p, X v) {
  typedef __attribute__((address_space(1))) X GX;
  *(GX*)p = v;
}

// BF16 storage of the FP32 phase's MDIIS residual history, which enters
// only the Gram products <r_i, r_j>: round to nearest even, four values in
// eight bytes.  The Gram row of a new residual is formed from its rounded
// values, so the Gram matrix stays that of the stored vectors.
struct alignas(8) BF4 { uint32_t lo, hi; };
__device__ __forceinline__ uint32_t bf16_bits(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ BF4 ldbf4(const uint16_t* p) {
  typedef __attribute__((address_space(1))) const u32x2 G;
  const u32x2 x = *(G*)p;
  return BF4{x[0], x[1]};
}
__device__ __forceinline__ void stbf4(uint16_t* p, BF4 q) {
  typedef __attribute__((address_space(1))) u32x2 G;
  u32x2 x;
  x[0] = q.lo;
  x[1] = q.hi;
  *(G*)p = x;
}
// FP8 e4m3 (OCP: largest finite 448; the conversion does not saturate, so
// values are clamped first) of four points scaled by s, packed in a word.
__device__ __forceinline__ uint32_t to_fp8x4(const Vec16<float>& v, float s) {
  auto c = [s](float x) { return fminf(fmaxf(x * s, -448.f), 448.f); };
  int w = 0;
  w = __builtin_amdgcn_cvt_pk_fp8_f32(c(v.v[0]), c(v.v[1]), w, false);
  w = __builtin_amdgcn_cvt_pk_fp8_f32(c(v.v[2]), c(v.v[3]), w, true);
  return (uint32_t)w;
}
__device__ __forceinline__ Vec16<float> from_fp8x4(uint32_t w, float inv) {
  Vec16<float> x;
  x.v[0] = __builtin_amdgcn_cvt_f32_fp8((int)w, 0) * inv;
  x.v[1] = __builtin_amdgcn_cvt_f32_fp8((int)w, 1) * inv;
  x.v[2] = __builtin_amdgcn_cvt_f32_fp8((int)w, 2) * inv;
  x.v[3] = __builtin_amdgcn_cvt_f32_fp8((int)w, 3) * inv;
  return x;
}
__device__ __forceinline__ BF4 to_bf4(const Vec16<float>& v) {
  BF4 q;
  q.lo = bf16_bits(v.v[0]) | (bf16_bits(v.v[1]) << 16);
  q.hi = bf16_bits(v.v[2]) | (bf16_bits(v.v[3]) << 16);
  return q;
}
// The FP32 phase's residual history (Gram rows only): four points of entry
// j at element idx as one packed word (BF16 pairs, or an FP8 word in lo),
// unpacked to FP32, and the new r stored rounded (returned as rounded, so
// that the Gram row is formed from the values the history keeps).
__device__ __forceinline__ BF4 ld_hist(const SoluteDesc& d, const void* p, int64_t idx) {
  if (d.r8) {
    BF4 q;
    q.lo = *reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(p) + idx);
    q.hi = 0;
    return q;
  }
  return ldbf4(static_cast<const uint16_t*>(p) + idx);
}
__device__ __forceinline__ Vec16<float> un_hist(const SoluteDesc& d, int j, BF4 q) {
  if (d.r8) return from_fp8x4(q.lo, d.rhinv[j]);
  Vec16<float> x;
  x.v[0] = bf16_val(q.lo & 0xFFFFu);
  x.v[1] = bf16_val(q.lo >> 16);
  x.v[2] = bf16_val(q.hi & 0xFFFFu);
  x.v[3] = bf16_val(q.hi >> 16);
  return x;
}
__device__ __forceinline__ Vec16<float> st_hist(const SoluteDesc& d, int64_t idx,
                                                const Vec16<float>& r) {
  if (d.r8) {
    const uint32_t w = to_fp8x4(r, d.rsc);
    *reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(d.r) + idx) = w;
    return from_fp8x4(w, 1.0f / d.rsc);
  }
  const BF4 q = to_bf4(r);
  stbf4(static_cast<uint16_t*>(d.r) + idx, q);
  Vec16<float> x;
  x.v[0] = bf16_val(q.lo & 0xFFFFu);
  x.v[1] = bf16_val(q.lo >> 16);
  x.v[2] = bf16_val(q.hi & 0xFFFFu);
  x.v[3] = bf16_val(q.hi >> 16);
  return x;
}

// Index of the k point (kz, ky, kx) in a solute's form factor F(k) s(k):
// z-major rows of NH like the tables in FP64; in FP32 y-major rows of
// N0/2 bins (the layout of the plane chain's spectra) and the Nyquist
// column kx = N0/2 compact behind them, so that the z pass reads it
// as it reads the spectra.
template <typename T>
__device__ __forceinline__ int64_t fs_index(const GridArgs& g, int z, int y, int x) {
  if constexpr (sizeof(T) == 8) return ((int64_t)z * g.N1 + y) * g.NH + x;
  return x == g.NX - 1 ? g.K + (int64_t)z * g.N1 + y : ((int64_t)y * g.N2 + z) * g.NH + x;
}

// One vector of VW points at e (an offset into the slab off of solute d's
// site arrays): the MDIIS step t = sum_j coef_j u_j (written to t) or the
// given t, then c = closure(beta v, t) in the solver mode of d.
template <typename T, int H>
__device__ __forceinline__ Vec16<T> closure_vec(const SoluteDesc& d, const GridArgs& g,
                                                int64_t off, int e, T* __restrict__ t,
                                                const T* __restrict__ bv) {
  constexpr int VW = Vec16<T>::W;
  const int nc = d.ncomb;
  Vec16<T> c;
  Vec16<T> tv;
  const Vec16<T> b = ld16(bv + e);
  if (nc > 0) {
    double acc[VW];
#pragma unroll
    for (int k = 0; k < VW; ++k) acc[k] = 0.0;
    bool summed = false;
    if constexpr (sizeof(T) == 4) {
      if (d.u16) {  // the newest u (FP32) and BF16 differences, loads first
        const Vec16<T> u0 = ld16(static_cast<const T*>(d.u[0]) + off + e);
        BF4 du[H];
#pragma unroll
        for (int j = 1; j < H; ++j)
          if (j < nc) du[j] = ldbf4(static_cast<const uint16_t*>(d.u[j]) + off + e);
#pragma unroll
        for (int k = 0; k < VW; ++k) acc[k] = d.coef[0] * (double)u0.v[k];
#pragma unroll
        for (int j = 1; j < H; ++j)
          if (j < nc) {
            const double cj = d.coef[j];
            acc[0] += cj * (double)bf16_val(du[j].lo & 0xFFFFu);
            acc[1] += cj * (double)bf16_val(du[j].lo >> 16);
            acc[2] += cj * (double)bf16_val(du[j].hi & 0xFFFFu);
            acc[3] += cj * (double)bf16_val(du[j].hi >> 16);
          }
        summed = true;
      }
    }
    if (!summed) {
      // all loads first, then the sums: independent loads in flight
      Vec16<T> u[H];
#pragma unroll
      for (int j = 0; j < H; ++j)
        if (j < nc) u[j] = ld16(static_cast<const T*>(d.u[j]) + off + e);
      Vec16<T> sh;
      if (d.shift_w != 0.0) sh = ld16(static_cast<const T*>(d.shift) + off + e);
#pragma unroll
      for (int j = 0; j < H; ++j)
        if (j < nc) {
          const double cj = d.coef[j];
#pragma unroll
          for (int k = 0; k < VW; ++k) acc[k] += cj * (double)u[j].v[k];
        }
      if (d.shift_w != 0.0) {
#pragma unroll
        for (int k = 0; k < VW; ++k) acc[k] -= d.shift_w * (double)sh.v[k];
      }
    }
#pragma unroll
    for (int k = 0; k < VW; ++k) tv.v[k] = T(acc[k]);
    st16(t + e, tv);
  } else {
    tv = ld16(t + e);
  }
#pragma unroll
  for (int k = 0; k < VW; ++k) {
    if (d.mode == 0) c.v[k] = closure_c<T>(g.closure, g.pse_order, b.v[k], tv.v[k]);
    else if (d.mode == 5)  // bv holds x_b (PY: exp(-beta v) - 1, c linear in t)
      c.v[k] = (g.closure == 2) ? b.v[k] * tv.v[k]
                                : closure_dc<T>(g.closure, g.pse_order, b.v[k], tv.v[k]);
    else if (d.mode == 6)  // J v: dc = (dc/dt) v (compute_c1, hnc3d.c:1307)
      c.v[k] = T(gld(d.dt + off + e + k) * (double)tv.v[k]);
    else if (d.mode == 7)  // flag form: dc + v (hnc3d.c:1310-1320)
      c.v[k] = T((gld(d.dt + off + e + k) + 1.0) * (double)tv.v[k]);
    else if (d.mode == 10)  // rho h for the susceptibility (hnc3d.c:1006-1030)
      c.v[k] = T(g.rho) * (closure_c<T>(g.closure, g.pse_order, b.v[k], tv.v[k]) + tv.v[k]);
    else c.v[k] = tv.v[k];
  }
  return c;
}

// ------------------------------------------- 1. closure fused with x r2c
// t = sum_j coef_j u_j (the MDIIS step; written back as the iterate) or the
// given t, then c = closure(beta v, t), then the x-axis r2c of c.
// Grid: x = ceil(N1 N2 / ROWS), y = site, z = solute.
template <typename T, int N0, int H>
__global__ void __launch_bounds__(NT)
closure_r2c_x(const SoluteDesc* __restrict__ sd, GridArgs g,
              const cpx<T>* __restrict__ tw_half, const cpx<T>* __restrict__ tw_full) {
  constexpr int M = N0 / 2;
  constexpr int ROWS = rows_for(M);
  constexpr int LSM = lstride_for(M);
  __shared__ cpx<T> lds[ROWS * LSM];
  const SoluteDesc& d = sd[blockIdx.z];
  if (!d.active) return;
  const int site = blockIdx.y;
  const int nrows = g.N1 * g.N2;
  const int row0 = blockIdx.x * ROWS;
  const int64_t off = site * g.P + (int64_t)row0 * N0;
  T* __restrict__ t = static_cast<T*>(d.t) + off;
  const T* __restrict__ bv = static_cast<const T*>(d.bv) + off;
  const int vrows = min(ROWS, nrows - row0);
  const int nvalid = vrows * N0;
  constexpr int VW = Vec16<T>::W;
  for (int e = threadIdx.x * VW; e < ROWS * N0; e += NT * VW) {
    Vec16<T> c;
#pragma unroll
    for (int k = 0; k < VW; ++k) c.v[k] = T(0);
    if (e < nvalid) c = closure_vec<T, H>(d, g, off, e, t, bv);  // whole vectors
    // z_n = c_{2n} + i c_{2n+1} of row e / N0
    cpx<T>* zrow = lds + (e / N0) * LSM;
    const int n0 = (e % N0) / 2;
#pragma unroll
    for (int h = 0; h < VW / 2; ++h) zrow[lpos(n0 + h)] = cpx<T>{c.v[2 * h], c.v[2 * h + 1]};
  }
  __syncthreads();
  fft_lines<T, M, ROWS, NT, false>(lds, LSM, vrows, tw_half);
  cpx<T>* W = static_cast<cpx<T>*>(d.W) + site * g.K + (int64_t)row0 * g.NH;
  const int nout = min(ROWS, nrows - row0) * (M + 1);
  for (int e = threadIdx.x; e < nout; e += NT) {
    const int row = e / (M + 1);
    const int k = e - row * (M + 1);
    const cpx<T> zk = lds[row * LSM + lpos(k == M ? 0 : k)];
    const cpx<T> zm = cconj(lds[row * LSM + lpos(k == 0 ? 0 : M - k)]);
    const cpx<T> s = cadd(zk, zm);
    const cpx<T> dd = rot90<T, false>(cmul(tw_full[k], csub(zk, zm)));  // -i w (Zk - conj Zm)
    stc(W + row * g.NH + k, cpx<T>{T(0.5) * (s.re + dd.re), T(0.5) * (s.im + dd.im)});
  }
}

// -------------------------------------------------- 2/4. c2c along y
// Grid: x = ceil(NH / LINES), y = site * N2 + z, z = solute.
template <typename T, int N1, int LINES, bool INV>
__global__ void __launch_bounds__(NT)
fft_y(const SoluteDesc* __restrict__ sd, GridArgs g, const cpx<T>* __restrict__ tw) {
  __shared__ cpx<T> lds[LINES * N1];  // interleaved: point y of column l at y * LINES + l
  const SoluteDesc& d = sd[blockIdx.z];
  if (!d.active) return;
  const int site = blockIdx.y / g.N2;
  const int z = blockIdx.y - site * g.N2;
  const int kx0 = blockIdx.x * LINES;
  cpx<T>* W = static_cast<cpx<T>*>(d.W) + site * g.K + (int64_t)z * g.N1 * g.NH + kx0;
  {
    // all loads first (the row padding keeps every column of a tile in
    // bounds; columns past NX are zeroed after the load)
    constexpr int PTS = N1 * LINES, PT = (PTS + NT - 1) / NT;
    const int nl = min(LINES, g.NX - kx0);
    cpx<T> v[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const int e = threadIdx.x + p * NT;
      if (PT * NT == PTS || e < PTS) {
        const int y = e / LINES, l = e - y * LINES;
        // the last tile (the Nyquist column) reads only its valid columns
        if (nl == LINES || l < nl) v[p] = ldc(W + y * g.NH + l);
      }
    }
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const int e = threadIdx.x + p * NT;
      if (PT * NT == PTS || e < PTS) {
        const int l = e & (LINES - 1);
        lds[e] = (kx0 + l < g.NX) ? v[p] : cpx<T>{T(0), T(0)};
      }
    }
  }
  __syncthreads();
  fft_lines_il<T, N1, LINES, NT, INV>(lds, min(LINES, g.NX - kx0), tw);
  for (int e = threadIdx.x; e < N1 * LINES; e += NT) {
    const int y = e / LINES, l = e - y * LINES;
    if (kx0 + l < g.NX) stc(W + y * g.NH + l, lds[e]);
  }
}

// |k|^2 = (kx^2 + ky^2) + kz^2, rounded exactly as the reference's
// SQR(k[0]) + SQR(k[1]) + SQR(k[2]) (bgy3d-vec.h:664-667): no FMA
// contraction.  interp() below is discontinuous at table midpoints and grid
// k often lands exactly on one, so the last bit selects the neighbour.
__device__ __forceinline__ double kmag2(double kx, double ky, double kz) {
#pragma clang fp contract(off)
  const double a = kx * kx;
  const double b = ky * ky;
  const double c = kz * kz;
  return (a + b) + c;
}

// 1D table lookup (interp, bgy3d-vec.c:23-46) at |k|; the index set is
// shared by every table at one k-point.
struct Interp {
  int i, ip, im;
  double off;  // ix - i
};
__device__ __forceinline__ Interp interp_at(double k, double dk, int n) {
  const double ix = k / dk - 0.5;
  int i = (int)(ix + 0.5);
  i = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);  // host guarantees range; clamp never faults
  Interp r;
  r.i = i;
  r.ip = (i < n - 2) ? i + 1 : i;
  r.im = (i > 1) ? i - 1 : i;
  r.off = ix - i;
  return r;
}
__device__ __forceinline__ double interp_eval(const Interp& a, const double* __restrict__ tab) {
  const double d = tab[a.ip] - tab[a.im];
  const double fi = (a.ip - a.im == 2) ? d * 0.5 : d;  // == d / (ip - im), exactly
  return tab[a.i] + fi * a.off;
}

// chi_ij(|k|) - delta_ij and tau_i(|k|) on the 3D k grid, once per batch
// (solute independent; solvent_kernel_rism hnc3d.c:1443-1566 tabulates
// them the same way with vec_ktab).  Grid: x = chunks of K.  Writes FP64
// and FP32 copies, [npairs][K] and [m][K].
__global__ void __launch_bounds__(NT)
ktab3(GridArgs g, const double* __restrict__ chi_tab, const double* __restrict__ tau_tab,
      double* __restrict__ chi3, float* __restrict__ chi3f, double* __restrict__ tau3,
      float* __restrict__ tau3f) {
  const int np = g.m * (g.m + 1) / 2;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.K; e += stride) {
    const int x = (int)(e % g.NH);
    if (x >= g.NX) continue;  // row padding
    const int y = (int)((e / g.NH) % g.N1);
    const int z = (int)(e / ((int64_t)g.NH * g.N1));
    const int iy = (y <= g.N1 / 2) ? y : y - g.N1;
    const int iz = (z <= g.N2 / 2) ? z : z - g.N2;
    const double k2 = kmag2(x * g.dkx, iy * g.dky, iz * g.dkz);
    const Interp ip = interp_at(sqrt(k2), g.dk_tab, g.nrad);
    for (int p = 0; p < np; ++p) {
      const double v = interp_eval(ip, chi_tab + (int64_t)p * g.nrad);
      chi3[p * g.K + e] = v;
      chi3f[p * g.K + e] = (float)v;
    }
    for (int i = 0; i < g.m; ++i) {
      const double v = interp_eval(ip, tau_tab + (int64_t)i * g.nrad);
      tau3[i * g.K + e] = v;
      tau3f[i * g.K + e] = (float)v;
    }
  }
}

// The per-k solvent-solvent OZ of one k point (iterate_t2, hnc3d.c:376-470;
// compute_t2_m, hnc3d.c:231-324) for g.msol <= MSX sites, the site-site
// pair spectra of the point in LDS (pair p = i(i+1)/2 + j at
// lds[(p LINES + l) LS + lpos(z)]), overwritten with T.  sgn is the corner
// phase of the point, U the unit long-range Coulomb there, kk its index
// into the omega table chi3.
template <typename T, int MSX, int LINES, int LS>
__device__ void oz_pairs(cpx<T>* __restrict__ lds, const GridArgs& g, const T* __restrict__ chi3,
                         int64_t kk, int l, int z, double sgn, double U) {
  // solvent-solvent OZ (iterate_t2, hnc3d.c:376-470; compute_t2_m,
  // hnc3d.c:231-324): C = h^3 c s - beta q_i q_j U at the corner,
  // X = (1 - rho W C)^-1 W C W, T = ((X - C) - beta q_i q_j U) s / L^3
  const int ms = g.msol;
  double Cr[MSX][MSX], Ci[MSX][MSX], Wm[MSX][MSX];
  for (int i = 0; i < ms; ++i)
    for (int j = 0; j <= i; ++j) {
      const int p = i * (i + 1) / 2 + j;
      const cpx<T> c = lds[(p * LINES + l) * LS + lpos(z)];
      const double q2 = g.q[i] * g.q[j];
      const double re = (double)c.re * g.h3 * sgn - g.beta * q2 * U;
      const double im = (double)c.im * g.h3 * sgn;
      Cr[i][j] = Cr[j][i] = re;
      Ci[i][j] = Ci[j][i] = im;
      Wm[i][j] = Wm[j][i] = (i == j) ? 1.0 : (double)chi3[p * g.K + kk];
    }
  // WC, A = 1 - rho WC, B = WC W
  double WCr[MSX][MSX], WCi[MSX][MSX], Ar[MSX][MSX], Ai[MSX][MSX], Br[MSX][MSX],
      Bi[MSX][MSX];
  for (int i = 0; i < ms; ++i)
    for (int j = 0; j < ms; ++j) {
      double ar = 0.0, ai = 0.0;
      for (int q = 0; q < ms; ++q) { ar += Wm[i][q] * Cr[q][j]; ai += Wm[i][q] * Ci[q][j]; }
      WCr[i][j] = ar;
      WCi[i][j] = ai;
      Ar[i][j] = (i == j ? 1.0 : 0.0) - g.rho * ar;
      Ai[i][j] = -g.rho * ai;
    }
  for (int i = 0; i < ms; ++i)
    for (int j = 0; j < ms; ++j) {
      double br = 0.0, bi = 0.0;
      for (int q = 0; q < ms; ++q) { br += WCr[i][q] * Wm[q][j]; bi += WCi[i][q] * Wm[q][j]; }
      Br[i][j] = br;
      Bi[i][j] = bi;
    }
  // Gaussian elimination with partial pivoting, A X = B
  for (int col = 0; col < ms; ++col) {
    int piv = col;
    double best = Ar[col][col] * Ar[col][col] + Ai[col][col] * Ai[col][col];
    for (int r = col + 1; r < ms; ++r) {
      const double a2 = Ar[r][col] * Ar[r][col] + Ai[r][col] * Ai[r][col];
      if (a2 > best) { best = a2; piv = r; }
    }
    if (piv != col)
      for (int q = 0; q < ms; ++q) {
        double tr = Ar[col][q], ti = Ai[col][q];
        Ar[col][q] = Ar[piv][q]; Ai[col][q] = Ai[piv][q]; Ar[piv][q] = tr; Ai[piv][q] = ti;
        tr = Br[col][q]; ti = Bi[col][q];
        Br[col][q] = Br[piv][q]; Bi[col][q] = Bi[piv][q]; Br[piv][q] = tr; Bi[piv][q] = ti;
      }
    const double dr = Ar[col][col], di = Ai[col][col];
    const double den = dr * dr + di * di;
    const double ir = dr / den, ii = -di / den;  // 1 / A[col][col]
    for (int r = col + 1; r < ms; ++r) {
      const double fr = Ar[r][col] * ir - Ai[r][col] * ii;
      const double fi = Ar[r][col] * ii + Ai[r][col] * ir;
      for (int q = col; q < ms; ++q) {
        Ar[r][q] -= fr * Ar[col][q] - fi * Ai[col][q];
        Ai[r][q] -= fr * Ai[col][q] + fi * Ar[col][q];
      }
      for (int q = 0; q < ms; ++q) {
        Br[r][q] -= fr * Br[col][q] - fi * Bi[col][q];
        Bi[r][q] -= fr * Bi[col][q] + fi * Br[col][q];
      }
    }
  }
  for (int col = ms - 1; col >= 0; --col) {
    const double dr = Ar[col][col], di = Ai[col][col];
    const double den = dr * dr + di * di;
    const double ir = dr / den, ii = -di / den;
    for (int q = 0; q < ms; ++q) {
      double xr = Br[col][q], xi = Bi[col][q];
      for (int r = col + 1; r < ms; ++r) {
        xr -= Ar[col][r] * Br[r][q] - Ai[col][r] * Bi[r][q];
        xi -= Ar[col][r] * Bi[r][q] + Ai[col][r] * Br[r][q];
      }
      Br[col][q] = xr * ir - xi * ii;
      Bi[col][q] = xr * ii + xi * ir;
    }
  }
  for (int i = 0; i < ms; ++i)
    for (int j = 0; j <= i; ++j) {
      const int p = i * (i + 1) / 2 + j;
      const double q2 = g.q[i] * g.q[j];
      const double tr = ((Br[i][j] - Cr[i][j]) - g.beta * q2 * U) * sgn * g.inv_L3;
      const double ti = (Bi[i][j] - Ci[i][j]) * sgn * g.inv_L3;
      lds[(p * LINES + l) * LS + lpos(z)] = cpx<T>{T(tr), T(ti)};
    }
}

// --------------------------- 3. z c2c + chi contraction + tau + z c2c^-1
// Grid: zorder 0: x = solute (concurrent workgroups share one chi3/tau3
// tile), y = ceil(NH / LINES), z = ky row; zorder 1: x = kx tile, y = ky
// row, z = solute (concurrent workgroups stream contiguous planes of one
// solute).  Dynamic LDS:
// m * LINES * lstride_for(N2) complex.  chi3: [m(m+1)/2][K] chi_ij - delta_ij;
// tau3: [m][K] s_i(|k|), both on the k grid, in the working type.
template <typename T, int N2, int LINES>
__global__ void __launch_bounds__(NT)
zstar(const SoluteDesc* __restrict__ sd, GridArgs g, const cpx<T>* __restrict__ tw,
      const T* __restrict__ chi3, const T* __restrict__ tau3) {
  constexpr int LS = lstride_for(N2);
  extern __shared__ unsigned char smem[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem);
  // zorder 0: grid (solute, kx tile, ky); 1: grid (kx tile, ky, solute)
  const int b = g.zorder ? blockIdx.z : blockIdx.x;
  const SoluteDesc& d = sd[b];
  if (!d.active) return;
  const int m = g.m;
  const int y = g.zorder ? blockIdx.y : blockIdx.z;
  const int kx0 = (g.zorder ? blockIdx.x : blockIdx.y) * LINES;
  const int64_t zs = (int64_t)g.N1 * g.NH;  // z stride
  const int64_t base = (int64_t)y * g.NH + kx0;
  cpx<T>* W = static_cast<cpx<T>*>(d.W) + base;
  for (int s = 0; s < m; ++s) {
    const cpx<T>* Ws = W + s * g.K;
    cpx<T>* L = lds + s * LINES * LS;
    for (int e = threadIdx.x; e < N2 * LINES; e += NT) {
      const int z = e / LINES, l = e - z * LINES;
      L[l * LS + lpos(z)] = (kx0 + l < g.NX) ? Ws[z * zs + l] : cpx<T>{T(0), T(0)};
    }
  }
  __syncthreads();
  const int nl = min(LINES, g.NX - kx0);  // valid kx columns of this tile
  if (d.mode != 4) {  // mode 4: inverse z FFTs of the loaded spectrum only
  for (int s = 0; s < m; ++s) fft_lines<T, N2, LINES, NT, false>(lds + s * LINES * LS, LS, nl, tw);
  if (d.mode == 3) {  // forward spectrum only
    for (int s = 0; s < m; ++s) {
      cpx<T>* Ws = W + s * g.K;
      const cpx<T>* L = lds + s * LINES * LS;
      for (int e = threadIdx.x; e < N2 * LINES; e += NT) {
        const int z = e / LINES, l = e - z * LINES;
        if (kx0 + l < g.NX) Ws[z * zs + l] = L[l * LS + lpos(z)];
      }
    }
    return;
  }
  const bool linear = d.mode != 0;  // chi convolution only

  const int iy = (y <= g.N1 / 2) ? y : y - g.N1;
  const double ky = iy * g.dky;
  if (g.special != 0) {
    for (int e = threadIdx.x; e < N2 * LINES; e += NT) {
      const int z = e / LINES, l = e - z * LINES;
      if (kx0 + l >= g.NX) continue;
      const int64_t kk = base + z * zs + l;
      const int iz = (z <= N2 / 2) ? z : z - N2;
      const int ix = kx0 + l;
      const double sgn = ((ix + iy + iz) & 1) ? -1.0 : 1.0;
      if (g.special == 2) {
        // Poisson: u(k) = 4 pi/eps0 h^3 rho(k) / k^2, 0 at k = 0, k^2 summed
        // z, y, x (bgy3d_poisson, bgy3d-poisson.c:103-190); / L^3 for the
        // inverse FFT
        const double kz = iz * g.dkz, kx = ix * g.dkx;
        const double k2 = (kz * kz + ky * ky) + kx * kx;
        const double fac = k2 != 0.0 ? 4 * M_PI * EPSILON0INV * g.h3 / k2 * g.inv_L3 : 0.0;
        for (int s = 0; s < m; ++s) {
          cpx<T>& v = lds[(s * LINES + l) * LS + lpos(z)];
          v = cscale(v, T(fac));
        }
        continue;
      }
      // unit long-range Coulomb as vec_kmap tabulates it: k = |k|, then
      // 4 pi/eps0 exp(-k^2/4G^2)/k^2 (bgy3d-force.c:19-36, bgy3d-force.h)
      const double kabs = sqrt(kmag2(ix * g.dkx, ky, iz * g.dkz));
      const double k2 = kabs * kabs;
      const double U = k2 == 0.0 ? 0.0
          : (4 * M_PI * EPSILON0INV) * exp(-k2 / (4 * G_COULOMB * G_COULOMB)) / k2;
      if (g.special == 3) {
        for (int s = 0; s < m; ++s)
          lds[(s * LINES + l) * LS + lpos(z)] = cpx<T>{T(U * sgn * g.inv_L3), T(0)};
        continue;
      }
      if (g.msol <= 3) oz_pairs<T, 3, LINES, LS>(lds, g, chi3, kk, l, z, sgn, U);
      else oz_pairs<T, MAX_SITES, LINES, LS>(lds, g, chi3, kk, l, z, sgn, U);
    }
  } else {
  const cpx<T>* Fs = static_cast<const cpx<T>*>(d.Fs);
  for (int e = threadIdx.x; e < N2 * LINES; e += NT) {
    const int z = e / LINES, l = e - z * LINES;
    if (kx0 + l >= g.NX) continue;
    const int64_t kk = base + z * zs + l;   // k-grid index of this point
    const cpx<T> fs = Fs[fs_index<T>(g, z, y, kx0 + l)];
    cpx<T> c[MAX_SITES];
    for (int s = 0; s < m; ++s) c[s] = lds[(s * LINES + l) * LS + lpos(z)];
    cpx<T> uc = {T(0), T(0)};
    if (!g.renorm && !linear) {
      // u_c(k) = 4 pi eps0^-1 exp(-k^2/4G^2)/k^2 F(k) s(k), 0 at k = 0
      // (bgy3d-force.h:140-147, bgy3d-solutes.c:767-792).
      const int iz = (z <= N2 / 2) ? z : z - N2;
      const double k2 = kmag2((kx0 + l) * g.dkx, ky, iz * g.dkz);
      const double shape = (k2 == 0.0) ? 0.0
          : (4 * M_PI * EPSILON0INV) * exp(-k2 / (4 * G_COULOMB * G_COULOMB)) / k2;
      uc = cscale(fs, T(shape));
      for (int s = 0; s < m; ++s) {
        // h^3 C_j - beta q_j u_c  (hnc3d.c:1226-1227)
        c[s] = csub(cscale(c[s], T(g.h3)), cscale(uc, T(g.beta * g.q[s])));
      }
    }
    for (int i = 0; i < m; ++i) {
      cpx<T> acc = {T(0), T(0)};
      for (int j = 0; j < m; ++j) {  // sum_j S_ij sig_j c_j / sig_i (sig = 1: one density)
        const int pi_ = (i >= j) ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
        const T chi = chi3[pi_ * g.K + kk] * T(g.sig[j]);
        acc.re += chi * c[j].re;
        acc.im += chi * c[j].im;
      }
      acc = cscale(acc, T(1.0 / g.sig[i]));
      cpx<T> out;
      if (linear) {
        out = cscale(acc, T(g.a_scale));
      } else if (g.renorm) {
        // (h^3/L^3) sum_j chi_ij C_j - (beta/eps0/L^3) tau_i F s  (hnc3d.c:1245-1268)
        const T tau = tau3[i * g.K + kk] * T(g.tau_scale);
        out = csub(cscale(acc, T(g.a_scale)), cscale(fs, tau));
      } else {
        out = cscale(csub(acc, cscale(uc, T(g.beta * g.q[i]))), T(g.inv_L3));
      }
      lds[(i * LINES + l) * LS + lpos(z)] = out;
    }
  }
  }  // g.special
  }  // d.mode != 4
  __syncthreads();
  for (int s = 0; s < m; ++s) fft_lines<T, N2, LINES, NT, true>(lds + s * LINES * LS, LS, nl, tw);
  for (int s = 0; s < m; ++s) {
    cpx<T>* Ws = W + s * g.K;
    const cpx<T>* L = lds + s * LINES * LS;
    for (int e = threadIdx.x; e < N2 * LINES; e += NT) {
      const int z = e / LINES, l = e - z * LINES;
      if (kx0 + l < g.NX) Ws[z * zs + l] = L[l * LS + lpos(z)];
    }
  }
}

// ------------------------------------------------ block reductions
__device__ __forceinline__ double warp_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double warp_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// --------------------------- 3b. zstar with the spectra held in registers
// Mode 3: forward z FFTs only; mode 4: inverse z FFTs only.
// The solute residual's k-space step (modes 0-7 of zstar, special 0) for a
// compile-time site count MS: the z lines of one site at a time pass
// through LDS (LINES * N2 points, interleaved), and each thread keeps its PPT
// points of all MS sites in registers for the contraction.  LDS per
// workgroup is 1/MS of zstar's, so a 128-byte row of kx columns (16 FP32
// or 8 FP64) fits at high occupancy.  Same arithmetic as zstar, point by
// point.  Grid as zstar.
template <typename T, int N2>
constexpr int zreg_lines() {
  int L = 128 / (int)sizeof(cpx<T>);
  const int pmax = sizeof(T) == 4 ? 8 : 4;  // points per thread per site
  while (L > 2 && (N2 * L + NT - 1) / NT > pmax) L /= 2;
  return L;
}

template <typename T, int N2, int LINES, int MS, bool NYQ, int NTZ = NT>
__global__ void __launch_bounds__(NTZ)
zstar_reg(const SoluteDesc* __restrict__ sd, GridArgs g, const cpx<T>* __restrict__ tw,
          const T* __restrict__ chi3, const T* __restrict__ tau3) {
  constexpr int PTS = N2 * LINES;
  constexpr int PPT = (PTS + NTZ - 1) / NTZ;
  __shared__ cpx<T> lds[PTS];  // interleaved: point z of line l at z * LINES + l
  // Block order.  zorder 2 (1D grid): consecutive blocks go to the XCDs in
  // turn, and each XCD walks all solutes of one (ky, kx tile) before the
  // next, so the tile's chi and tau stay in that XCD's L2 for every solute.
  // zorder 3 (1D grid): XCD x owns the rows ky = x (mod 8) and walks every
  // solute of a row, the row's kx tiles back to back, so each z line's
  // columns are read by neighbouring workgroups at once (whole rows of the
  // z planes instead of one tile's 128 bytes) and the row's chi and tau
  // stay in the XCD's L2 across solutes.
  // zorder 1: x = kx tile, y = ky, z = solute; zorder 0: x = solute,
  // y = kx tile, z = ky.
  int b, tile, ky;
  if (NYQ) {
    b = blockIdx.z;
    tile = 0;
    ky = blockIdx.x * LINES;
  } else if (g.zorder == 3) {
    const int lin = blockIdx.x, xcd = lin & 7, slot = lin >> 3;
    tile = slot % g.ztiles;
    b = (slot / g.ztiles) % g.zB;
    ky = (slot / (g.ztiles * g.zB)) * 8 + xcd;
  } else if (g.zorder == 2) {
    const int lin = blockIdx.x, xcd = lin & 7, slot = lin >> 3;
    const int gt = (slot / g.zB) * 8 + xcd;  // global tile, ky-major
    b = slot - (slot / g.zB) * g.zB;
    tile = gt % g.ztiles;
    ky = gt / g.ztiles;
  } else if (g.zorder == 1) {
    b = blockIdx.z;
    tile = blockIdx.x;
    ky = blockIdx.y;
  } else {
    b = blockIdx.x;
    tile = blockIdx.y;
    ky = blockIdx.z;
  }
  const SoluteDesc& d = sd[b];
  if (!d.active) return;
  // NYQ: lines are the rows y0 + l of the Nyquist column kx = N0/2 (grid
  // x = ceil(N1 / LINES), y = 1, z = solute); otherwise the kx columns
  // kx0 + l of row y.
  const int y = ky;
  const int kx0 = NYQ ? g.NX - 1 : tile * LINES;
  // element offsets below 2^32 bytes from per-workgroup bases (at32): the
  // tables and Fs from their starts, W from its site's
  const int zs = g.N1 * g.NH;         // z stride (Fs, chi, tau)
  const int ls = NYQ ? g.NH : 1;      // line stride
  const int base = y * g.NH + kx0;
  // W: the same index unless ylay (y-major rows, compact Nyquist column)
  int wz = zs, wl = ls;
  int64_t wsite = g.K;
  cpx<T>* W = static_cast<cpx<T>*>(d.W) + base;
  if (g.ylay) {
    if (NYQ) {
      W = static_cast<cpx<T>*>(d.W) + g.m * g.K + y;
      wz = g.N1;
      wl = 1;
      wsite = (int64_t)g.N1 * g.N2;
    } else {
      W = static_cast<cpx<T>*>(d.W) + (int64_t)y * g.N2 * g.NH + kx0;
      wz = g.NH;
    }
  }
  const int nl = NYQ ? min(LINES, g.N1 - y) : min(LINES, g.NX - kx0);  // valid lines
  const int tid = threadIdx.x;
  cpx<T> c[MS][PPT];
  // every site's loads first, staged in its registers of c (its FFT result
  // lands there after its pass), so the later sites' loads overlap the
  // earlier sites' transforms; a partial tile reads only its valid lines
#pragma unroll
  for (int s = 0; s < MS; ++s) {
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int e = tid + p * NTZ;
      if (PPT * NTZ == PTS || e < PTS) {
        const int z = e / LINES, l = e - z * LINES;
        c[s][p] = (nl == LINES || l < nl) ? ldc(at32(W + s * wsite, (uint32_t)(z * wz + l * wl)))
                                          : cpx<T>{T(0), T(0)};
      }
    }
  }
  // RR: the z transforms keep the thread's points in registers (its PPT
  // points of line l are the inputs of one first-stage butterfly and the
  // outputs of its last-stage butterflies, fft_lines_rr); otherwise they
  // pass through LDS whole (fft_lines_il).
  constexpr bool RR = rr_ok(N2, LINES, NTZ) && PPT * NTZ == PTS;
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    cpx<T>* Ws = W + s * wsite;
    if constexpr (RR) {
      if (d.mode == 4) continue;  // inverse only: the loaded spectrum goes to the inverse pass
      fft_lines_rr<T, N2, LINES, NTZ, false>(c[s], lds, nl, tw);
      if (d.mode == 3) {  // forward spectrum only
#pragma unroll
        for (int p = 0; p < PPT; ++p) {
          const int e = tid + p * NTZ;
          const int z = e / LINES, l = e - z * LINES;
          if (l < nl) stc(at32(Ws, (uint32_t)(z * wz + l * wl)), c[s][p]);
        }
      }
      continue;
    }
    {
      cpx<T>* v = c[s];
      if (d.mode == 4) continue;  // inverse only: the loaded spectrum goes to the inverse pass
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const int e = tid + p * NTZ;
        if (PPT * NTZ == PTS || e < PTS) lds[e] = ((e & (LINES - 1)) < nl) ? v[p] : cpx<T>{T(0), T(0)};
      }
    }
    __syncthreads();
    fft_lines_il<T, N2, LINES, NTZ, false>(lds, nl, tw);
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int e = tid + p * NTZ;
      if (PPT * NTZ == PTS || e < PTS) {
        const int z = e / LINES, l = e - z * LINES;
        c[s][p] = lds[e];
        if (d.mode == 3 && l < nl) stc(at32(Ws, (uint32_t)(z * wz + l * wl)), c[s][p]);  // forward spectrum only
      }
    }
    __syncthreads();
  }
  if (d.mode == 3) return;
  const bool linear = d.mode != 0;  // chi convolution only
  const cpx<T>* Fs = static_cast<const cpx<T>*>(d.Fs);
#pragma unroll
  for (int p = 0; p < PPT; ++p) {
    const int e = tid + p * NTZ;
    if (d.mode == 4 || !(PPT * NTZ == PTS || e < PTS)) continue;
    const int z = e / LINES, l = e - z * LINES;
    if (l >= nl) continue;
    const uint32_t kk = (uint32_t)(base + z * zs + l * ls);  // k-grid index of this point
    cpx<T> fs = {T(0), T(0)};
    if (!linear) fs = ldc(at32(Fs, (uint32_t)fs_index<T>(g, z, NYQ ? y + l : y, kx0 + (NYQ ? 0 : l))));
    cpx<T> uc = {T(0), T(0)};
    if (!g.renorm && !linear) {
      // u_c(k) = 4 pi eps0^-1 exp(-k^2/4G^2)/k^2 F(k) s(k), 0 at k = 0
      // (bgy3d-force.h:140-147, bgy3d-solutes.c:767-792).
      const int iz = (z <= N2 / 2) ? z : z - N2;
      const int yl = NYQ ? y + l : y;
      const int iy = (yl <= g.N1 / 2) ? yl : yl - g.N1;
      const double k2 = kmag2((NYQ ? kx0 : kx0 + l) * g.dkx, iy * g.dky, iz * g.dkz);
      const double shape = (k2 == 0.0) ? 0.0
          : (4 * M_PI * EPSILON0INV) * exp(-k2 / (4 * G_COULOMB * G_COULOMB)) / k2;
      uc = cscale(fs, T(shape));
#pragma unroll
      for (int s = 0; s < MS; ++s)  // h^3 C_j - beta q_j u_c  (hnc3d.c:1226-1227)
        c[s][p] = csub(cscale(c[s][p], T(g.h3)), cscale(uc, T(g.beta * g.q[s])));
    }
    T chi[MS * (MS + 1) / 2];
#pragma unroll
    for (int q = 0; q < MS * (MS + 1) / 2; ++q) chi[q] = gld(at32(chi3 + q * g.K, kk));
    cpx<T> out[MS];
#pragma unroll
    for (int i = 0; i < MS; ++i) {
      cpx<T> acc = {T(0), T(0)};
#pragma unroll
      for (int j = 0; j < MS; ++j) {  // sum_j S_ij sig_j c_j / sig_i (sig = 1: one density)
        const T x = chi[(i >= j) ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i] * T(g.sig[j]);
        acc.re += x * c[j][p].re;
        acc.im += x * c[j][p].im;
      }
      acc = cscale(acc, T(1.0 / g.sig[i]));
      if (linear) {
        out[i] = cscale(acc, T(g.a_scale));
      } else if (g.renorm) {
        // (h^3/L^3) sum_j chi_ij C_j - (beta/eps0/L^3) tau_i F s  (hnc3d.c:1245-1268)
        const T tau = gld(at32(tau3 + i * g.K, kk)) * T(g.tau_scale);
        out[i] = csub(cscale(acc, T(g.a_scale)), cscale(fs, tau));
      } else {
        out[i] = cscale(csub(acc, cscale(uc, T(g.beta * g.q[i]))), T(g.inv_L3));
      }
    }
#pragma unroll
    for (int i = 0; i < MS; ++i) c[i][p] = out[i];
  }
#pragma unroll
  for (int s = 0; s < MS; ++s) {
    cpx<T>* Ws = W + s * wsite;
    if constexpr (RR) {
      fft_lines_rr<T, N2, LINES, NTZ, true>(c[s], lds, nl, tw);
#pragma unroll
      for (int p = 0; p < PPT; ++p) {
        const int e = tid + p * NTZ;
        const int z = e / LINES, l = e - z * LINES;
        if (l < nl) stc(at32(Ws, (uint32_t)(z * wz + l * wl)), c[s][p]);
      }
      continue;
    }
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int e = tid + p * NTZ;
      if (PPT * NTZ == PTS || e < PTS) lds[e] = c[s][p];
    }
    __syncthreads();
    fft_lines_il<T, N2, LINES, NTZ, true>(lds, nl, tw);
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
      const int e = tid + p * NTZ;
      if (PPT * NTZ == PTS || e < PTS) {
        const int z = e / LINES, l = e - z * LINES;
        if (l < nl) stc(at32(Ws, (uint32_t)(z * wz + l * wl)), lds[e]);
      }
    }
    __syncthreads();
  }
}

// -------------------------------------- 5. x c2r fused with the residual
// t' = c2r(...), r = t' - t, u = t + lam r, and per workgroup the partial
// max|r|, sum r^2 and dots <r, r_j> with the history (MDIIS Gram row), to
// part[(site * gridDim.x + blockIdx.x) * NPART + q].  Grid as closure_r2c_x.
template <typename T, int N0, int H>
__global__ void __launch_bounds__(NT)
c2r_x_resid(const SoluteDesc* __restrict__ sd, GridArgs g,
            const cpx<T>* __restrict__ tw_half, const cpx<T>* __restrict__ tw_full) {
  constexpr int M = N0 / 2;
  constexpr int ROWS = rows_for(M);
  __shared__ cpx<T> xb[ROWS * (M + 1)];
  constexpr int LSM = lstride_for(M);
  __shared__ cpx<T> zb[ROWS * LSM];
  __shared__ double red[NPART][NT / 64];
  const SoluteDesc& d = sd[blockIdx.z];
  if (!d.active) return;
  const int site = blockIdx.y;
  const int nrows = g.N1 * g.N2;
  const int row0 = blockIdx.x * ROWS;
  const int vrows = min(ROWS, nrows - row0);
  const cpx<T>* W = static_cast<const cpx<T>*>(d.W) + site * g.K + (int64_t)row0 * g.NH;
  const int64_t off = site * g.P + (int64_t)row0 * N0;
  constexpr int VW = Vec16<T>::W;
  // FP32 defect-correction passes over whole tiles: the streams this
  // thread reads after the inverse FFT (t, F(t_b), the BF16 residual
  // history) are loaded together with the spectrum, so those loads overlap
  // the FFT instead of following it.  ITP vectors per thread.
  constexpr int ITP = (ROWS * N0) / (NT * VW);
  constexpr bool PREF = sizeof(T) == 4 && ITP * NT * VW == ROWS * N0 && ITP <= 2;
  const bool pref = PREF && (d.r16 != 0 || d.r8 != 0) && d.mode == 5 && vrows == ROWS;
  Vec16<T> ptv[PREF ? ITP : 1], pfb[PREF ? ITP : 1];
  BF4 prh[PREF ? ITP : 1][H];
  {
    // all loads first; rows past the last one read the last row's address
    // and are zeroed
    constexpr int PTS = ROWS * (M + 1), PT = (PTS + NT - 1) / NT;
    cpx<T> v[PT];
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const int e = threadIdx.x + p * NT;
      if (PT * NT == PTS || e < PTS) {
        const int row = e / (M + 1);
        const int k = e - row * (M + 1);
        v[p] = ldc(W + min(row, vrows - 1) * g.NH + k);
      }
    }
    if constexpr (PREF) {
      if (pref) {
#pragma unroll
        for (int it = 0; it < ITP; ++it) {
          const int64_t e = off + threadIdx.x * VW + it * NT * VW;
          ptv[it] = ld16(static_cast<const T*>(d.t) + e);
          pfb[it] = ld16(static_cast<const T*>(d.fb) + e);
#pragma unroll
          for (int j = 0; j < H; ++j)
            if (j < d.ndot) prh[it][j] = ld_hist(d, d.rh[j], e);
        }
      }
    }
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const int e = threadIdx.x + p * NT;
      if (PT * NT == PTS || e < PTS) xb[e] = (e / (M + 1) < vrows) ? v[p] : cpx<T>{T(0), T(0)};
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ROWS * M; e += NT) {
    const int row = e / M;
    const int k = e - row * M;
    cpx<T> xk = xb[row * (M + 1) + k];
    cpx<T> xm = xb[row * (M + 1) + (M - k)];
    if (k == 0) { xk.im = T(0); xm.im = T(0); }  // c2r ignores Im at DC / Nyquist
    const cpx<T> a = cadd(xk, cconj(xm));
    const cpx<T> b = csub(xk, cconj(xm));
    const cpx<T> ib = rot90<T, true>(cmulc(b, tw_full[k]));  // i e^{+2 pi i k/N} b
    zb[row * LSM + lpos(k)] = cadd(a, ib);
  }
  __syncthreads();
  fft_lines<T, M, ROWS, NT, true>(zb, LSM, vrows, tw_half);
  const T* __restrict__ t = static_cast<const T*>(d.t) + off;
  T* __restrict__ r = static_cast<T*>(d.r) + off;  // (BF16 slots: addressed in e below)
  T* __restrict__ un = d.unew ? static_cast<T*>(d.unew) + off : nullptr;  // null: no u
  const bool r16 = d.r16 != 0 || d.r8 != 0;  // FP32 phase: r history in BF16 or FP8
  const int nd = d.ndot;
  const T lam = T(d.lam);
  double mx = 0.0, ss = 0.0;
  double dot[H];
#pragma unroll
  for (int j = 0; j < H; ++j) dot[j] = 0.0;
  const int mode = d.mode;
  // one vector of VW points at e: t, F(t_b) (mode 5) and the history rh
  auto point = [&](int e, const Vec16<T>& tv, const Vec16<T>& f, const Vec16<T>* rh) {
    Vec16<T> z;  // t' at points e.. of row e / N0: z_n = t'_{2n} + i t'_{2n+1}
    {
      const cpx<T>* zrow = zb + (e / N0) * LSM;
      const int n0 = (e % N0) / 2;
#pragma unroll
      for (int h = 0; h < VW / 2; ++h) {
        const cpx<T> zz = zrow[lpos(n0 + h)];
        z.v[2 * h] = zz.re;
        z.v[2 * h + 1] = zz.im;
      }
    }
    Vec16<T> rv, uv;
    if (mode == 5) {  // defect correction: F(t_b) + K dc - delta
#pragma unroll
      for (int k = 0; k < VW; ++k) rv.v[k] = f.v[k] + z.v[k] - tv.v[k];
    } else {
#pragma unroll
      for (int k = 0; k < VW; ++k) {
        if (mode == 0 || mode == 6) rv.v[k] = z.v[k] - tv.v[k];                     // F(t), J v
        else if (mode == 1)                                                          // adjoint
          rv.v[k] = T(gld(d.dt + off + e + k)) * z.v[k] + T(gld(d.gv + off + e + k)) - tv.v[k];
        else rv.v[k] = z.v[k];                                                       // K t
      }
    }
#pragma unroll
    for (int k = 0; k < VW; ++k) uv.v[k] = tv.v[k] + lam * rv.v[k];
    if (un) {
      if constexpr (sizeof(T) == 4) {
        if (d.u16) {  // the difference from the u it replaces, in BF16
          const Vec16<T> uo = ld16(un + e);
          Vec16<T> du;
#pragma unroll
          for (int k = 0; k < VW; ++k) du.v[k] = uv.v[k] - uo.v[k];
          stbf4(static_cast<uint16_t*>(d.dnew) + off + e, to_bf4(du));
        }
      }
      st16(un + e, uv);
    }
#pragma unroll
    for (int k = 0; k < VW; ++k) mx = fmax(mx, fabs((double)rv.v[k]));
    bool stored = false;
    if constexpr (sizeof(T) == 4) {
      if (r16) {  // store rounded, and form the Gram row from the rounded values
        rv = st_hist(d, off + e, rv);
        stored = true;
      }
    }
    if (!stored) st16(r + e, rv);
    // products of one vector in the working type, sums across vectors in FP64
    {
      T s2 = T(0);
#pragma unroll
      for (int k = 0; k < VW; ++k) s2 += rv.v[k] * rv.v[k];
      ss += (double)s2;
#pragma unroll
      for (int j = 0; j < H; ++j)
        if (j < nd) {
          T pj = T(0);
#pragma unroll
          for (int k = 0; k < VW; ++k) pj += rv.v[k] * rh[j].v[k];
          dot[j] += (double)pj;
        }
    }
  };
  auto unbf = [&](int j, BF4 q) {
    Vec16<T> x;
    if constexpr (sizeof(T) == 4) x = un_hist(d, j, q);
    return x;
  };
  bool done = false;
  if constexpr (PREF) {
    if (pref) {
#pragma unroll
      for (int it = 0; it < ITP; ++it) {
        Vec16<T> rh[H];
#pragma unroll
        for (int j = 0; j < H; ++j)
          if (j < nd) rh[j] = unbf(j, prh[it][j]);
        point(threadIdx.x * VW + it * NT * VW, ptv[it], pfb[it], rh);
      }
      done = true;
    }
  }
  if (!done) {
    for (int e = threadIdx.x * VW; e < vrows * N0; e += NT * VW) {
      const Vec16<T> tv = ld16(t + e);
      Vec16<T> f;  // F(t_b), defect correction
      if (mode == 5) f = ld16(static_cast<const T*>(d.fb) + off + e);
      Vec16<T> rh[H];
#pragma unroll
      for (int j = 0; j < H; ++j)
        if (j < nd) {
          if constexpr (sizeof(T) == 4) {
            if (r16) {
              rh[j] = unbf(j, ld_hist(d, d.rh[j], off + e));
              continue;
            }
          }
          rh[j] = ld16(static_cast<const T*>(d.rh[j]) + off + e);
        }
      point(e, tv, f, rh);
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  mx = warp_max(mx);
  ss = warp_sum(ss);
  if (lane == 0) { red[0][w] = mx; red[1][w] = ss; }
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < nd) {
      const double s = warp_sum(dot[j]);
      if (lane == 0) red[2 + j][w] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 + nd) {
    const int q = threadIdx.x;
    double a = red[q][0];
    for (int i = 1; i < NT / 64; ++i) a = (q == 0) ? fmax(a, red[q][i]) : a + red[q][i];
    gst(d.part + ((int64_t)site * gridDim.x + blockIdx.x) * NPART + q, a);
  }
}

// ------------------------------- 1' and 5'. the xy-plane pair (FP32)
// The FP32 residual in three launches: closure_r2c_xy, zstar, c2r_xy_resid.
// A workgroup owns one z plane of one site of one solute: N1 rows of N0
// points, whose half spectrum (N1 rows of N0/2 + 1 bins, 66.5 KB in FP32
// at 128^2) stays in LDS between the x and the y transforms, so the y pass
// of the five-launch chain costs no HBM traffic.  Rows of the plane are at
// stride LS = N0/2 + 1 in LDS; the row FFTs run on the first N0/2 points of
// each row, the column FFTs on all N0/2 + 1 columns.
constexpr int NTP = 512;  // threads of a plane workgroup

// t, c = closure(beta v, t) as closure_r2c_x over the plane, the N0-point
// r2c of every row (N0/2-point FFT of z_n = c_2n + i c_2n+1 and the split,
// in place: bins k and N0/2 - k from Z_k and Z_{N0/2-k}), then the c2c of
// every kx column along y; the plane's rows [y][kx] go to W.
// Grid: x = z plane (N2), y = site, z = solute.
template <typename T, int N0, int N1, int H>
__global__ void __launch_bounds__(NTP)
closure_r2c_xy(const SoluteDesc* __restrict__ sd, GridArgs g, const cpx<T>* __restrict__ tw_half,
               const cpx<T>* __restrict__ tw_full, const cpx<T>* __restrict__ tw_y) {
  constexpr int M = N0 / 2, LS = M + 1, PLANE = N0 * N1, HP = M / 2 + 1;
  constexpr int VW = Vec16<T>::W;
  static_assert(PLANE % VW == 0, "whole vectors");
  constexpr bool WHOLE = PLANE % (2 * NTP * VW) == 0;  // batches of two vectors
  __shared__ cpx<T> lds[N1 * LS];
  const SoluteDesc& d = sd[blockIdx.z];
  if (!d.active) return;
  const int site = blockIdx.y, z = blockIdx.x;
  const int64_t off = site * g.P + (int64_t)z * PLANE;
  T* __restrict__ t = static_cast<T*>(d.t) + off;
  const T* __restrict__ bv = static_cast<const T*>(d.bv) + off;
  // two vectors per step, so that both vectors' streams are in flight together
  for (int e0 = threadIdx.x * VW; e0 < PLANE; e0 += 2 * NTP * VW) {
    Vec16<T> c[2];
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (WHOLE || e0 + q * NTP * VW < PLANE)
        c[q] = closure_vec<T, H>(d, g, off, e0 + q * NTP * VW, t, bv);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = e0 + q * NTP * VW;
      if (!WHOLE && e >= PLANE) continue;
      cpx<T>* zrow = lds + (e / N0) * LS;
      const int n0 = (e % N0) / 2;
#pragma unroll
      for (int h = 0; h < VW / 2; ++h) zrow[n0 + h] = cpx<T>{c[q].v[2 * h], c[q].v[2 * h + 1]};
    }
  }
  __syncthreads();
  fft_lines<T, M, N1, NTP, false>(lds, LS, N1, tw_half);
  for (int e = threadIdx.x; e < N1 * HP; e += NTP) {
    const int row = e / HP, k = e - row * HP;
    cpx<T>* r = lds + row * LS;
    const cpx<T> a = r[k], b = r[k == 0 ? 0 : M - k];  // Z_k, Z_{M-k}
    // bin kk from Z_kk and Z_{M-kk}: (Z + conj Z') / 2 - i w_kk (Z - conj Z') / 2
    auto bin = [&](int kk, cpx<T> zk, cpx<T> zo) {
      const cpx<T> zm = cconj(zo);
      const cpx<T> sm = cadd(zk, zm);
      const cpx<T> dd = rot90<T, false>(cmul(tw_full[kk], csub(zk, zm)));
      return cpx<T>{T(0.5) * (sm.re + dd.re), T(0.5) * (sm.im + dd.im)};
    };
    if (k == 0) {
      r[0] = bin(0, a, a);
      r[M] = bin(M, a, a);
    } else {
      const cpx<T> xk = bin(k, a, b), xm = bin(M - k, b, a);
      r[k] = xk;
      r[M - k] = xm;
    }
  }
  __syncthreads();
  fft_cols<T, N1, LS, M + 1, NTP, false>(lds, tw_y);
  // y-major rows of M bins, then the Nyquist bins of the plane's rows
  cpx<T>* W = static_cast<cpx<T>*>(d.W) + site * g.K + (int64_t)z * g.NH;
  const int64_t ys = (int64_t)g.N2 * g.NH;
  for (int e = threadIdx.x; e < N1 * M; e += NTP) {
    const int y = e / M, k = e - y * M;
    stc(W + y * ys + k, lds[y * LS + k]);
  }
  cpx<T>* Q = static_cast<cpx<T>*>(d.W) + g.m * g.K + ((int64_t)site * g.N2 + z) * N1;
  for (int y = threadIdx.x; y < N1; y += NTP) stc(Q + y, lds[y * LS + M]);
}

// The inverse: the plane's rows [y][kx] from W, the inverse c2c of every
// column along y, the c2r pre-processing of every row in place (Z_k and
// Z_{N0/2-k} from bins k and N0/2 - k; Im of the DC and Nyquist bins
// dropped, as FFTW's c2r does), the inverse N0/2-point row FFTs, then
// r = t' - t (or the mode's form), u = t + lam r and the partials max|r|,
// sum r^2, <r, r_j> as c2r_x_resid, one partial row per plane at
// part[(site * N2 + z) * NPART + q].  Grid as closure_r2c_xy.
template <typename T, int N0, int N1, int H>
__global__ void __launch_bounds__(NTP)
c2r_xy_resid(const SoluteDesc* __restrict__ sd, GridArgs g, const cpx<T>* __restrict__ tw_half,
             const cpx<T>* __restrict__ tw_full, const cpx<T>* __restrict__ tw_y) {
  constexpr int M = N0 / 2, LS = M + 1, PLANE = N0 * N1, HP = M / 2 + 1;
  constexpr int VW = Vec16<T>::W;
  static_assert(PLANE % VW == 0, "whole vectors");
  constexpr bool WHOLE = PLANE % (2 * NTP * VW) == 0;  // batches of two vectors
  __shared__ cpx<T> lds[N1 * LS];
  __shared__ double red[NPART][NTP / 64];
  const SoluteDesc& d = sd[blockIdx.z];
  if (!d.active) return;
  const int site = blockIdx.y, z = blockIdx.x;
  // the plane's y-major rows of M bins and its Nyquist bins (closure_r2c_xy)
  const cpx<T>* W = static_cast<const cpx<T>*>(d.W) + site * g.K + (int64_t)z * g.NH;
  const cpx<T>* Q = static_cast<const cpx<T>*>(d.W) + g.m * g.K + ((int64_t)site * g.N2 + z) * N1;
  const int64_t ys = (int64_t)g.N2 * g.NH;
  {
    constexpr int PTS = N1 * M, PT = (PTS + NTP - 1) / NTP;
    constexpr int QT = (N1 + NTP - 1) / NTP;
    cpx<T> v[PT], nq[QT];
#pragma unroll
    for (int p = 0; p < PT; ++p) {  // all loads first
      const int e = threadIdx.x + p * NTP;
      if (PT * NTP == PTS || e < PTS) {
        const int y = e / M, k = e - y * M;
        v[p] = ldc(W + y * ys + k);
      }
    }
#pragma unroll
    for (int p = 0; p < QT; ++p) {
      const int y = threadIdx.x + p * NTP;
      if (y < N1) nq[p] = ldc(Q + y);
    }
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const int e = threadIdx.x + p * NTP;
      if (PT * NTP == PTS || e < PTS) {
        const int y = e / M, k = e - y * M;
        lds[y * LS + k] = v[p];
      }
    }
#pragma unroll
    for (int p = 0; p < QT; ++p) {
      const int y = threadIdx.x + p * NTP;
      if (y < N1) lds[y * LS + M] = nq[p];
    }
  }
  __syncthreads();
  fft_cols<T, N1, LS, M + 1, NTP, true>(lds, tw_y);
  for (int e = threadIdx.x; e < N1 * HP; e += NTP) {
    const int row = e / HP, k = e - row * HP;
    cpx<T>* r = lds + row * LS;
    cpx<T> xk = r[k], xm = r[M - k];  // bins k and M - k (0: DC and Nyquist)
    // Z_kk = (X_kk + conj X_{M-kk}) + i e^{+2 pi i kk/N0} (X_kk - conj X_{M-kk})
    auto pre = [&](int kk, cpx<T> a0, cpx<T> b0) {
      const cpx<T> a = cadd(a0, cconj(b0));
      const cpx<T> b = csub(a0, cconj(b0));
      return cadd(a, rot90<T, true>(cmulc(b, tw_full[kk])));
    };
    if (k == 0) {
      xk.im = T(0);
      xm.im = T(0);
      r[0] = pre(0, xk, xm);
    } else {
      const cpx<T> zk = pre(k, xk, xm), zm = pre(M - k, xm, xk);
      r[k] = zk;
      r[M - k] = zm;
    }
  }
  __syncthreads();
  fft_lines<T, M, N1, NTP, true>(lds, LS, N1, tw_half);
  const int64_t off = site * g.P + (int64_t)z * PLANE;
  const T* __restrict__ t = static_cast<const T*>(d.t) + off;
  T* __restrict__ r = static_cast<T*>(d.r) + off;  // (BF16 slots: addressed in e below)
  T* __restrict__ un = d.unew ? static_cast<T*>(d.unew) + off : nullptr;  // null: no u
  const bool r16 = d.r16 != 0 || d.r8 != 0;  // FP32 phase: r history in BF16 or FP8
  const int nd = d.ndot;
  const int mode = d.mode;
  const T lam = T(d.lam);
  double mx = 0.0, ss = 0.0;
  double dot[H];
#pragma unroll
  for (int j = 0; j < H; ++j) dot[j] = 0.0;
  // one vector of VW points at e: t' from LDS, the streams t, F(t_b)
  // (mode 5) and the history r_j (rhj(j), unpacked on use)
  auto point = [&](int e, const Vec16<T>& tv, const Vec16<T>& f, auto&& rhj) {
    Vec16<T> uo;  // u16: the u the new one replaces (loaded first, used last)
    if constexpr (sizeof(T) == 4)
      if (d.u16 && un) uo = ld16(un + e);
    Vec16<T> zv;  // t' at points e.. of row e / N0
    {
      const cpx<T>* zrow = lds + (e / N0) * LS;
      const int n0 = (e % N0) / 2;
#pragma unroll
      for (int h = 0; h < VW / 2; ++h) {
        const cpx<T> zz = zrow[n0 + h];
        zv.v[2 * h] = zz.re;
        zv.v[2 * h + 1] = zz.im;
      }
    }
    Vec16<T> rv, uv;
#pragma unroll
    for (int k = 0; k < VW; ++k) {
      if (mode == 5) rv.v[k] = f.v[k] + zv.v[k] - tv.v[k];  // F(t_b) + K dc - delta
      else if (mode == 0 || mode == 6) rv.v[k] = zv.v[k] - tv.v[k];
      else if (mode == 1)
        rv.v[k] = T(gld(d.dt + off + e + k)) * zv.v[k] + T(gld(d.gv + off + e + k)) - tv.v[k];
      else rv.v[k] = zv.v[k];
    }
#pragma unroll
    for (int k = 0; k < VW; ++k) uv.v[k] = tv.v[k] + lam * rv.v[k];
    if (un) {
      if constexpr (sizeof(T) == 4) {
        if (d.u16) {  // the difference from the u it replaces (uo), in BF16
          Vec16<T> du;
#pragma unroll
          for (int k = 0; k < VW; ++k) du.v[k] = uv.v[k] - uo.v[k];
          stbf4(static_cast<uint16_t*>(d.dnew) + off + e, to_bf4(du));
        }
      }
      st16(un + e, uv);
    }
#pragma unroll
    for (int k = 0; k < VW; ++k) mx = fmax(mx, fabs((double)rv.v[k]));
    bool stored = false;
    if constexpr (sizeof(T) == 4) {
      if (r16) {  // store rounded, and form the Gram row from the rounded values
        rv = st_hist(d, off + e, rv);
        stored = true;
      }
    }
    if (!stored) st16(r + e, rv);
    T s2 = T(0);
#pragma unroll
    for (int k = 0; k < VW; ++k) s2 += rv.v[k] * rv.v[k];
    ss += (double)s2;
#pragma unroll
    for (int j = 0; j < H; ++j)
      if (j < nd) {
        const Vec16<T> x = rhj(j);
        T pj = T(0);
#pragma unroll
        for (int k = 0; k < VW; ++k) pj += rv.v[k] * x.v[k];
        dot[j] += (double)pj;
      }
  };
  bool done = false;
  if constexpr (sizeof(T) == 4) {
    if (r16) {
      // BF16 history (the FP32 phase): two vectors per step, their streams
      // in flight together, the history kept packed (two VGPRs per vector)
      for (int e0 = threadIdx.x * VW; e0 < PLANE; e0 += 2 * NTP * VW) {
        Vec16<T> tv[2], f[2];
        BF4 rb[2][H];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int e = e0 + q * NTP * VW;
          if (!WHOLE && e >= PLANE) continue;
          tv[q] = ld16(t + e);
          if (mode == 5) f[q] = ld16(static_cast<const T*>(d.fb) + off + e);
#pragma unroll
          for (int j = 0; j < H; ++j)
            if (j < nd) rb[q][j] = ld_hist(d, d.rh[j], off + e);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int e = e0 + q * NTP * VW;
          if (!WHOLE && e >= PLANE) continue;
          point(e, tv[q], f[q], [&](int j) { return un_hist(d, j, rb[q][j]); });
        }
      }
      done = true;
    }
  }
  if (!done) {
    for (int e = threadIdx.x * VW; e < PLANE; e += NTP * VW) {
      const Vec16<T> tv = ld16(t + e);
      Vec16<T> f;
      if (mode == 5) f = ld16(static_cast<const T*>(d.fb) + off + e);
      Vec16<T> rh[H];
#pragma unroll
      for (int j = 0; j < H; ++j)
        if (j < nd) rh[j] = ld16(static_cast<const T*>(d.rh[j]) + off + e);
      point(e, tv, f, [&](int j) { return rh[j]; });
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  mx = warp_max(mx);
  ss = warp_sum(ss);
  if (lane == 0) { red[0][w] = mx; red[1][w] = ss; }
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j < nd) {
      const double s2 = warp_sum(dot[j]);
      if (lane == 0) red[2 + j][w] = s2;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 + nd) {
    const int q = threadIdx.x;
    double a = red[q][0];
    for (int i = 1; i < NTP / 64; ++i) a = (q == 0) ? fmax(a, red[q][i]) : a + red[q][i];
    gst(d.part + ((int64_t)site * gridDim.x + blockIdx.x) * NPART + q, a);
  }
}

// Fixed-order reduction of the residual-workgroup partials of each solute.
// Grid: x = solute.  out[b * NPART + q]: max|r|, ||r||^2, <r, r_j>.
__global__ void __launch_bounds__(NT)
reduce_resid(const SoluteDesc* __restrict__ sd, int nparts, double* __restrict__ out) {
  __shared__ double red[NPART][NT / 64];
  const SoluteDesc& d = sd[blockIdx.x];
  if (!d.active) return;
  const int nq = 2 + d.ndot;
  double acc[NPART];
#pragma unroll
  for (int q = 0; q < NPART; ++q) acc[q] = 0.0;
  for (int i = threadIdx.x; i < nparts; i += NT) {
    const double* p = d.part + (int64_t)i * NPART;
    acc[0] = fmax(acc[0], p[0]);
#pragma unroll
    for (int q = 1; q < NPART; ++q)
      if (q < nq) acc[q] += p[q];
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  acc[0] = warp_max(acc[0]);
  if (lane == 0) red[0][w] = acc[0];
#pragma unroll
  for (int q = 1; q < NPART; ++q) {
    if (q < nq) {
      const double s = warp_sum(acc[q]);
      if (lane == 0) red[q][w] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < nq) {
    const int q = threadIdx.x;
    double a = red[q][0];
    for (int i = 1; i < NT / 64; ++i) a = (q == 0) ? fmax(a, red[q][i]) : a + red[q][i];
    out[blockIdx.x * NPART + q] = a;  // fmax drops NaNs; the sums keep them
  }
}

// v from lane `lane` (the same in every lane of the wavefront) to all lanes.
__device__ __forceinline__ double lane_bcast(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// MDIIS step of the FP32 defect-correction phase on the device, after the
// norm reduction of a pass (out): one wavefront per solute (grid x =
// solute).  For every solute of the pass it does what the host loop does
// between two passes: the stopping and rebase tests, the new entry into
// the history and its Gram row, the restart test, the DIIS coefficients,
// and the descriptor of the next pass (the combination t = u_newest -
// sum_i (c_0 + .. + c_i) (u_{i+1} - u_i) over the BF16 differences, the
// slot of the new entry, the FP8 scales).  A solute that wants a rebase or
// failed is left inactive for the host; with hold, one whose residual fell
// 1e3 below its base's goes on and is reported ready, and one that reached
// the phase's tolerance is left inactive as parked (the host groups
// rebases).  The DIIS system is solved as
// diis_coefficients solves it on the host, to the bit: LU with partial
// pivoting (first largest pivot), one step of refinement against the
// unfactored matrix, substitutions by columns in the host's order of
// summation, no contraction into fused multiply-adds.
__global__ void __launch_bounds__(64)
mdiis_step(const double* __restrict__ out, MdiisDev* __restrict__ mst,
           SoluteDesc* __restrict__ sd, MdiisStatus* __restrict__ status) {
#pragma clang fp contract(off)
  constexpr int NM = MAX_DIIS + 1;
  constexpr int MD = MAX_DIIS;
  __shared__ MdiisDev s;
  __shared__ double o[NPART];
  __shared__ double pcol[NM];  // |column| by row position, for the pivot scan
  __shared__ int plane[NM];    // lane holding the row at each position
  __shared__ int ctl[4];       // flag, -, -, LU ok
  const int b = blockIdx.x, l = threadIdx.x;
  SoluteDesc& d = sd[b];
  {  // every load at once: the state, the pass's norms and dots
    const int act = d.active;
    const uint64_t* src = reinterpret_cast<const uint64_t*>(mst + b);
    uint64_t* dst = reinterpret_cast<uint64_t*>(&s);
    for (int i = l; i < (int)(sizeof(MdiisDev) / 8); i += 64) dst[i] = src[i];
    if (l < NPART) o[l] = out[(size_t)b * NPART + l];
    if (!act) return;
  }
  __syncthreads();
  const double rinf = o[0], r2 = o[1];
  const bool fin = __builtin_isfinite(rinf) && __builtin_isfinite(r2);
  const int sn = s.slot_new, nd0 = s.ndot;
  if (fin && l < nd0) {  // the new entry after the ones it was dotted with, its Gram row
    const int q = s.dotted[l];
    s.order[l] = q;
    s.gram[sn * MD + q] = o[2 + l];
    s.gram[q * MD + sn] = o[2 + l];
  }
  if (l == 0) {
    s.iters++;
    s.iters32++;
    s.resid = rinf;
    int flag = 0;
    if (!fin) {
      flag = 2;
    } else {
      if (rinf < 0.7 * s.best32) { s.best32 = rinf; s.stall32 = 0; }
      else if (rinf <= (s.only32 ? 1e-3 : 1e-2 * s.rb_rinf)) { s.stall32++; }
      // with hold: the 1e3 drop only marks the solute ready, and one that
      // reached the phase's tolerance waits without iterating (parked)
      const bool drop = !s.only32 && rinf <= 1e-3 * s.rb_rinf, conv = rinf <= s.tol_half;
      if (s.stall32 >= 12 || s.iters32 >= s.max_iter || (!s.hold && (drop || conv)) ||
          (s.only32 && conv))
        flag = 1;
      else if (conv)
        flag = 5;
      else if (drop)
        flag = 4;
      s.sbase[sn] = s.base;
      s.order[nd0] = sn;
      s.n = nd0 + 1;
      s.gram[sn * MD + sn] = r2;
      // restart when the residual blows up against the best one, or stalls
      if (r2 < s.best) { s.best = r2; s.since_best = 0; } else { s.since_best++; }
      if (r2 > 100.0 * s.best || s.since_best > 3 * s.D) {
        s.order[0] = sn;
        s.n = 1;
        s.best = r2;
        s.since_best = 0;
      }
    }
    ctl[0] = flag;
    ctl[3] = 0;
  }
  __syncthreads();
  const int flag = ctl[0];
  if (flag != 2 && s.n > 1) {
    // bordered system [G/sc -1; -1 0] [c; l] = [0; -1], sc = max diag G
    const int nh = s.n, n = nh + 1;
    double sc = 0.0;
    for (int i = 0; i < nh; ++i) {
      const double a = fabs(s.gram[s.order[i] * MD + s.order[i]]);
      sc = sc < a ? a : sc;
    }
    if (sc > 0.0 && __builtin_isfinite(sc)) {
      // lane l < n holds row l of the matrix, a (factored in place) and a0
      // (as built); phys is the position the host's row swaps give that
      // row, so position i holds the row of lane la[i] and perm[i] = la[i]
      double a[NM], a0[NM];
      const int oi = l < nh ? s.order[l] : 0;
#pragma unroll
      for (int j = 0; j < NM; ++j) {
        double v = 0.0;
        if (l < n && j < n)
          v = (l < nh && j < nh) ? s.gram[oi * MD + s.order[j]] / sc
                                 : (l == nh && j == nh ? 0.0 : -1.0);
        a[j] = v;
        a0[j] = v;
      }
      int phys = l < n ? l : NM + l;
      bool ok = true;
#pragma unroll
      for (int col = 0; col < NM; ++col) {  // PA = LU, L unit lower
        if (!ok || col >= n) continue;
        // the host's scan over positions col..n-1: the first largest |A|
        if (l < n) {
          pcol[phys] = fabs(a[col]);
          plane[phys] = l;
        }
        __syncthreads();
        double v[NM];
        int pl[NM];
#pragma unroll
        for (int q = col; q < NM; ++q) {
          v[q] = pcol[q];
          pl[q] = plane[q];
        }
        __syncthreads();
        int piv = col, lp = pl[col];
        double pmag = v[col];
#pragma unroll
        for (int q = col + 1; q < NM; ++q)
          if (q < n && v[q] > pmag) {
            pmag = v[q];
            piv = q;
            lp = pl[q];
          }
        const int lc = pl[col];
        lp = __builtin_amdgcn_readfirstlane(lp);
        if (!(pmag > 1e-300)) { ok = false; continue; }
        if (l == lp) phys = col;  // positions col and piv swap rows
        else if (l == lc) phys = piv;
        const double inv = 1.0 / lane_bcast(a[col], lp);
        double pr[NM];
#pragma unroll
        for (int k = col + 1; k < NM; ++k) pr[k] = lane_bcast(a[k], lp);
        if (phys > col && l < n) {
          const double f = a[col] * inv;
          a[col] = f;
#pragma unroll
          for (int k = col + 1; k < NM; ++k)
            if (k < n) a[k] -= f * pr[k];
        }
      }
      if (ok) {
        int la[NM];
#pragma unroll
        for (int k = 0; k < NM; ++k) {
          const unsigned long long m = __ballot(phys == k);
          la[k] = __builtin_amdgcn_readfirstlane(m ? __ffsll(m) - 1 : 0);
        }
        // x = LU^-1 b with b_l in lane l, x_l to lane l, twice: b, then
        // the residual b - A x of one refinement step (x += A^-1 (b - A x)).
        // Row sums run by columns, in the host's order of summation.
        const double b0 = l == nh ? -1.0 : 0.0;
        double rhs = b0, xv = 0.0, c = 0.0;
#pragma unroll
        for (int it = 0; it < 2; ++it) {
          double t = rhs, y = 0.0;
#pragma unroll
          for (int k = 0; k < NM; ++k) {  // forward, L unit lower
            if (k < n) {
              const double yk = lane_bcast(t, la[k]);
              if (phys == k) y = t;
              if (phys > k && l < n) t -= a[k] * yk;
            }
          }
          t = y;
          double x = 0.0;
#pragma unroll
          for (int k = NM - 1; k >= 0; --k) {  // backward, the last column first
            if (k < n) {
              const double xk = lane_bcast(t, la[k]) / lane_bcast(a[k], la[k]);
              if (l == k) x = xk;
              if (phys < k && l < n) t -= a[k] * xk;
            }
          }
          if (it == 0) {
            xv = x;
            double r = b0;
#pragma unroll
            for (int k = 0; k < NM; ++k)
              if (k < n) {
                const double xk = lane_bcast(xv, k);
                r -= a0[k] * xk;
              }
            rhs = r;
          } else {
            c = xv + x;
          }
        }
        if (l < nh) s.coef[l] = c;
        const bool bad = __ballot(l < nh && !__builtin_isfinite(c)) != 0;
        if (l == 0) ctl[3] = bad ? 0 : 1;
      }
    }
  }
  __syncthreads();
  if (flag != 2 && (s.n == 1 || !ctl[3]) && l == 0) {  // one entry, or no usable solution: restart
    s.order[0] = s.slot_new;
    s.n = 1;
    s.coef[0] = 1.0;
  }
  __syncthreads();
  if (flag == 0 || flag == 4) {
    // the next pass: t = u_newest - sum_i (c_0 + .. + c_i) (u_{i+1} - u_i),
    // the new r and u in a free slot, else in the oldest one (dropped)
    const int n = s.n, D = s.D;
    bool used = l >= D;
    for (int i = 0; i < n; ++i) used |= s.order[i] == l;
    const unsigned long long busy = __ballot(used);
    const bool fresh = n < D;
    const int q = fresh ? __ffsll(~busy) - 1 : s.order[0];
    const int nd = fresh ? n : n - 1;
    if (l + 1 < n) {
      double cum = 0.0;
      for (int i = 0; i <= l; ++i) cum += s.coef[i];
      d.u[1 + l] = s.U + (size_t)s.order[l + 1] * s.mP;
      d.coef[1 + l] = -cum;
    }
    const float rsc = fp8_scale(rinf);
    if (l < nd) {
      const int qq = s.order[fresh ? l : l + 1];
      s.dotted[l] = qq;
      d.rh[l] = s.R + (size_t)qq * s.mP;
      d.rhinv[l] = 1.0f / s.rsc[qq];
    }
    if (l == 0) {
      d.ncomb = n;
      d.u[0] = s.ucur;
      d.coef[0] = 1.0;
      s.slot_new = q;
      s.ndot = nd;
      d.ndot = nd;
      s.rsc[q] = rsc;
      d.rsc = rsc;
      d.r = s.R + (size_t)q * s.mP;
      d.dnew = s.U + (size_t)q * s.mP;
    }
  } else if (l == 0) {
    d.active = 0;
  }
  __syncthreads();
  if (l == 0) {
    // the iteration count last, released to the host that polls for it
    MdiisStatus& st = status[b];
    st.flag = flag;
    st.iters32 = s.iters32;
    st.hist = (flag == 0 || flag == 4) ? (s.n > s.ndot ? s.n : s.ndot) : 0;
    st.resid = rinf;
    st.r2 = r2;
    __threadfence_system();
    __hip_atomic_store(&st.iters, s.iters, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&s);
    uint64_t* dst = reinterpret_cast<uint64_t*>(mst + b);
    for (int i = l; i < (int)(sizeof(MdiisDev) / 8); i += 64) dst[i] = src[i];
  }
}

// Host-prepared rows into the device MDIIS state and descriptors, read
// from pinned host memory: row r (grid x) is the solute index (8 bytes),
// its SoluteDesc, then its MdiisDev.
__global__ void __launch_bounds__(64)
mdiis_load(const uint64_t* __restrict__ rows, SoluteDesc* __restrict__ sd,
           MdiisDev* __restrict__ mst) {
  constexpr int WS = sizeof(SoluteDesc) / 8, WM = sizeof(MdiisDev) / 8;
  const uint64_t* row = rows + (size_t)blockIdx.x * (1 + WS + WM);
  const int b = (int)row[0];
  uint64_t* ds = reinterpret_cast<uint64_t*>(sd + b);
  uint64_t* dm = reinterpret_cast<uint64_t*>(mst + b);
  for (int i = threadIdx.x; i < WS; i += 64) ds[i] = row[1 + i];
  for (int i = threadIdx.x; i < WM; i += 64) dm[i] = row[1 + WS + i];
}

// FP32 <-> FP64 copies of the iterate at a precision hand-over.
// Grid: x = chunks, y = solute.
__global__ void __launch_bounds__(NT)
convert_batch(void* const* __restrict__ dst, const void* const* __restrict__ src, int to64,
              int64_t n, const int* __restrict__ active) {
  if (active && !active[blockIdx.y]) return;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride) {
    if (to64)
      static_cast<double*>(dst[blockIdx.y])[e] = (double)static_cast<const float*>(src[blockIdx.y])[e];
    else
      static_cast<float*>(dst[blockIdx.y])[e] = (float)static_cast<const double*>(src[blockIdx.y])[e];
  }
}

// u = t + lam r, FP64, for one solute (first MDIIS entry of an FP64 phase).
__global__ void __launch_bounds__(NT)
seed_u(double* __restrict__ u, const double* __restrict__ t, const double* __restrict__ r,
       double lam, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride)
    u[e] = t[e] + lam * r[e];
}

// Defect correction, base update: t_b += delta for one solute.
__global__ void __launch_bounds__(NT)
add_delta(double* __restrict__ tb, const float* __restrict__ delta, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride)
    tb[e] += (double)delta[e];
}

// Defect correction, the newest u of a difference-form history moved to
// the coordinates of a new base: u -= delta (the delta folded into t_b).
__global__ void __launch_bounds__(NT)
shift_u(float* __restrict__ u, const float* __restrict__ delta, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride) u[e] -= delta[e];
}

// Defect correction, new base t_b with F(t_b) = f64 known: x_b = -beta v +
// t_b (PY: exp(-beta v) - 1) into xb, F(t_b) into fb; with seed (the first
// base of a solve) also delta = 0 and the first MDIIS entry r_0 = F(t_b)
// (BF16 with r16), u_0 = 0 + lam r_0.
__global__ void __launch_bounds__(NT)
rebase_prep(const double* __restrict__ tb, const double* __restrict__ bv,
            const double* __restrict__ f64, int py, int seed, float lam,
            float* __restrict__ xb, float* __restrict__ fb, float* __restrict__ delta,
            float* __restrict__ u0, void* __restrict__ r0, int r16, float r8s, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride) {
    const double b = bv[e];
    xb[e] = py ? (float)expm1(-b) : (float)(-b + tb[e]);
    const float f = (float)f64[e];
    fb[e] = f;
    if (seed) {
      delta[e] = 0.f;
      if (r8s > 0.f)  // FP8 e4m3 of f * r8s (clamped: the conversion does not saturate)
        static_cast<uint8_t*>(r0)[e] = (uint8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(
            fminf(fmaxf(f * r8s, -448.f), 448.f), 0.f, 0, false) & 0xFF);
      else if (r16)
        static_cast<uint16_t*>(r0)[e] = (uint16_t)bf16_bits(f);
      else
        static_cast<float*>(r0)[e] = f;
      u0[e] = lam * f;
    }
  }
}

// ------------------------------------------------------ setup kernels
// Capped LJ and short-range Coulomb (bgy3d-solutes.c:45-131, 196-286).
__device__ __forceinline__ double lj0(double r) {
  const double p2 = 1.0 / (r * r);
  const double p6 = p2 * p2 * p2;
  return 4 * p6 * (p6 - 1);
}
__device__ __forceinline__ double lj1(double r) {
  const double p2 = 1.0 / (r * r);
  const double p6 = p2 * p2 * p2;
  return 4 * p6 * (6 - 12 * p6) / r;
}
__device__ __forceinline__ double ljcap0(double r) {
  const double R = 0.2;
  if (r >= R) return lj0(r);
  const double b = lj1(R) / (2 * R), a = lj0(R);
  return a + b * (r - R) * (r + R);
}
__device__ __forceinline__ double cs0(double r) { return erfc(r) / r; }
__device__ __forceinline__ double cs1(double r) {
  return -(erfc(r) + (2 / sqrt(M_PI)) * r * exp(-r * r)) / (r * r);
}
__device__ __forceinline__ double cscap0(double r) {
  const double R = 0.2;
  if (r >= R) return cs0(r);
  const double b = cs1(R) / (2 * R), a = cs0(R);
  return a + b * (r - R) * (r + R);
}
// erf(r)/r with its Taylor form below 0.02 (bgy3d-solutes.c:148-159)
__device__ __forceinline__ double cl0(double r) {
  if (r >= 0.02) return erf(r) / r;
  const double r2 = r * r;
  return (2 / sqrt(M_PI)) * (1. + r2 * (-1. / 3. + r2 * (1. / 10. + r2 * (-1. / 42.))));
}

// beta v_i(r) for every solvent site (field0, bgy3d-solutes.c:344-377).
// Grid: x = chunks of P, y = solute.  Writes FP64 and, if bv32 != NULL,
// an FP32 copy for the mixed-precision bulk iterations.  The solute sites
// pass through LDS in chunks of FB_CHUNK (position, sigma, sqrt(epsilon),
// charge / eps0); per point and site one r^2, 1 / r^2, r and the capped
// erfc(G r) / r (within G r < 6.5), then per solvent site the Lorentz-Berthelot pair
// (epsilon_ai = sqrt(eps_a) sqrt(eps_i), sigma_ai = (sigma_a + sigma_i) / 2)
// with (sigma_ai / r)^6 from 1 / r^2: ljcap0(r / sigma_ai) without a
// division or square root in the pair loop.
constexpr int FB_CHUNK = 256;
constexpr double FB_RC2 = (6.5 / G_COULOMB) * (6.5 / G_COULOMB);
// MS: the solvent-site arrays' size (>= g.m), so that water's three fit
// in few registers.
template <int MS>
__global__ void __launch_bounds__(NT)
field_bv(const SitesDesc* __restrict__ sites, GridArgs g, const double* __restrict__ solvent_sse,
         double L0, double L1, double L2, double h0, double h1, double h2,
         double* const* __restrict__ bv64, float* const* __restrict__ bv32) {
  __shared__ double st[FB_CHUNK][6];
  const SitesDesc sd = sites[blockIdx.y];
  const int m = g.m;
  double ssig[MS], sse[MS], sq[MS];
  for (int i = 0; i < m; ++i) {
    ssig[i] = solvent_sse[3 * i];
    sse[i] = sqrt(solvent_sse[3 * i + 1]);
    sq[i] = solvent_sse[3 * i + 2];
  }
  // ljcap0 below r = 0.2 sigma: a + b (x^2 - 0.04), x = r / sigma
  const double R = 0.2, cap_a = lj0(R), cap_b = lj1(R) / (2 * R);
  const int nch = (sd.n + FB_CHUNK - 1) / FB_CHUNK;
  auto load = [&](int c) {
    for (int k = threadIdx.x; k < FB_CHUNK; k += NT) {
      const int a = c * FB_CHUNK + k;
      if (a >= sd.n) break;
      const double* s = sd.xyz_s_e_q + 6 * a;
      st[k][0] = s[0];
      st[k][1] = s[1];
      st[k][2] = s[2];
      st[k][3] = s[3];
      st[k][4] = sqrt(s[4]);
      st[k][5] = s[5] * EPSILON0INV;
    }
  };
  if (nch == 1) {
    load(0);
    __syncthreads();
  }
  const int64_t stride = (int64_t)gridDim.x * NT;
  // rounds of NT points, the same count in every thread (LDS reloads)
  for (int64_t e0 = blockIdx.x * (int64_t)NT; e0 < g.P; e0 += stride) {
    const int64_t e = e0 + threadIdx.x;
    const bool ok = e < g.P;
    const int x = (int)(e % g.N0);
    const int y = (int)((e / g.N0) % g.N1);
    const int z = (int)(e / ((int64_t)g.N0 * g.N1));
    const double rx = x * h0 - L0 / 2, ry = y * h1 - L1 / 2, rz = z * h2 - L2 / 2;
    double v[MS];
    for (int i = 0; i < m; ++i) v[i] = 0.0;
    for (int c = 0; c < nch; ++c) {
      if (nch > 1) {
        __syncthreads();
        load(c);
        __syncthreads();
      }
      const int na = min(FB_CHUNK, sd.n - c * FB_CHUNK);
      if (!ok) continue;
      for (int k = 0; k < na; ++k) {
        const double dx = rx - st[k][0], dy = ry - st[k][1], dz = rz - st[k][2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double ir2 = 1.0 / r2;
        // erfc(G r) / r vanishes beyond G r = 6.5 (erfc(6.5) = 3.8e-20: the
        // term is below 2e-18 kcal/mol there, under the rounding of v)
        const double csc =
            r2 < FB_RC2 ? st[k][5] * G_COULOMB * cscap0(G_COULOMB * sqrt(r2)) : 0.0;
        for (int i = 0; i < m; ++i) {
          const double sab = 0.5 * (st[k][3] + ssig[i]);
          const double sab2 = sab * sab;
          const double s2 = sab2 * ir2;  // (sigma_ai / r)^2
          double lj;
          if (s2 <= 25.0) {
            const double p6 = s2 * s2 * s2;
            lj = 4 * p6 * (p6 - 1);
          } else {
            lj = cap_a + cap_b * (r2 / sab2 - R * R);
          }
          v[i] += st[k][4] * sse[i] * lj + sq[i] * csc;
        }
      }
    }
    if (!ok) continue;
    for (int i = 0; i < m; ++i) {
      const double b = g.beta * v[i];
      bv64[blockIdx.y][i * g.P + e] = b;
      if (bv32) bv32[blockIdx.y][i * g.P + e] = (float)b;
    }
  }
}

// Trigonometric interpolation of one solute's spectrum from the grid
// N/2 to N in the same box: Wf(k) = Wc(k) scale at every frequency the
// coarse grid holds, zero above.  The coarse Nyquist planes of the full y
// and z axes go in halves to +N_c/2 and -N_c/2, that of the half x axis in
// half to kx = N0_c/2 (the c2r mirrors it to -N0_c/2).  scale = 1 / P_c
// with the unnormalised inverse of the fine pipeline.  Grid: x = chunks
// of Kf, y = site.
__device__ __forceinline__ int up_index(int kf, int nf, int nc, double& w) {
  const int h = nc / 2;
  w = 1.0;
  if (kf < h) return kf;
  if (kf == h || kf == nf - h) { w = 0.5; return h; }
  if (kf > nf - h) return kf - nf + nc;
  return -1;
}
__global__ void __launch_bounds__(NT)
spec_upsample(const cpx<double>* __restrict__ wc, cpx<double>* __restrict__ wf, GridArgs g,
              int N0c, int N1c, int N2c, int NHc, int64_t Kc, double scale) {
  const int s = blockIdx.y;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.K; e += stride) {
    const int kx = (int)(e % g.NH);
    const int ky = (int)((e / g.NH) % g.N1);
    const int kz = (int)(e / ((int64_t)g.NH * g.N1));
    cpx<double> v = {0.0, 0.0};
    double wx = 1.0, wy, wz;
    const int hx = N0c / 2;
    const int cx = kx < hx ? kx : (kx == hx ? (wx = 0.5, hx) : -1);
    const int cy = up_index(ky, g.N1, N1c, wy);
    const int cz = up_index(kz, g.N2, N2c, wz);
    if (cx >= 0 && cy >= 0 && cz >= 0) {
      const cpx<double> c = wc[s * Kc + ((int64_t)cz * N1c + cy) * NHc + cx];
      const double w = wx * wy * wz * scale;
      v = cpx<double>{c.re * w, c.im * w};
    }
    wf[s * g.K + e] = v;
  }
}

// F(k) s(k) = s(k) sum_a Q_a exp(-i k.x_a), corner k (form_factor,
// bgy3d-solutes.c:587-600, times bgy3d_vec_fft_trans bgy3d-vec.c:619-646).
// Grid: x = chunks of K, y = solute.
__global__ void __launch_bounds__(NT)
form_factor(const SitesDesc* __restrict__ sites, GridArgs g, cpx<double>* const* __restrict__ fs64,
            cpx<float>* const* __restrict__ fs32) {
  const SitesDesc sd = sites[blockIdx.y];
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.K; e += stride) {
    const int x = (int)(e % g.NH);
    if (x >= g.NX) continue;  // row padding
    const int y = (int)((e / g.NH) % g.N1);
    const int z = (int)(e / ((int64_t)g.NH * g.N1));
    const int iy = (y <= g.N1 / 2) ? y : y - g.N1;
    const int iz = (z <= g.N2 / 2) ? z : z - g.N2;
    const double kx = x * g.dkx, ky = iy * g.dky, kz = iz * g.dkz;
    double re = 0.0, im = 0.0;
    for (int a = 0; a < sd.n; ++a) {
      const double* s = sd.xyz_s_e_q + 6 * a;
      double sn, cs;
      sincos(kx * s[0] + ky * s[1] + kz * s[2], &sn, &cs);
      re += s[5] * cs;
      im -= s[5] * sn;
    }
    const double sign = ((x + iy + iz) & 1) ? -1.0 : 1.0;
    fs64[blockIdx.y][e] = cpx<double>{sign * re, sign * im};
    if (fs32)
      fs32[blockIdx.y][fs_index<float>(g, z, y, x)] =
          cpx<float>{(float)(sign * re), (float)(sign * im)};
  }
}

// -------------------------------------------------- mu and observables
// Per point: c = closure(beta v, t), h = c + t, x = -beta v + t,
// cl_i = -beta q_i u_c(r) with u_c = eps0^-1 sum_a Q_a G cl0(G r)
// (coulomb_long, bgy3d-solutes.c:796-815); accumulate the chemical
// potential densities of the native closure, HNC, KH, GF and sum c
// (hnc3d.c:2172-2268), with Neumaier-compensated FP64 sums.  Writes
// g = h + 1 if gout != NULL.
// Grid: x = chunks of P, y = solute.  part[(b*gridDim.x + x)*5 + q].
__global__ void __launch_bounds__(NT)
mu_density(const SitesDesc* __restrict__ sites, GridArgs g, const double* const* __restrict__ tt,
           const double* const* __restrict__ bv64, double L0, double L1, double L2, double h0,
           double h1, double h2, double* const* __restrict__ gout, double* __restrict__ part) {
  __shared__ double red[5][NT / 64];
  const SitesDesc sd = sites[blockIdx.y];
  const double* t = tt[blockIdx.y];
  const double* bv = bv64[blockIdx.y];
  double* go = gout ? gout[blockIdx.y] : nullptr;
  double sum[5] = {0, 0, 0, 0, 0}, comp[5] = {0, 0, 0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.P; e += stride) {
    const int x = (int)(e % g.N0);
    const int y = (int)((e / g.N0) % g.N1);
    const int z = (int)(e / ((int64_t)g.N0 * g.N1));
    const double rx = x * h0 - L0 / 2, ry = y * h1 - L1 / 2, rz = z * h2 - L2 / 2;
    double uc = 0.0;
    for (int a = 0; a < sd.n; ++a) {
      const double* s = sd.xyz_s_e_q + 6 * a;
      const double dx = rx - s[0], dy = ry - s[1], dz = rz - s[2];
      // erf(x) rounds to 1 in FP64 beyond x = 6 (erfc(6) = 2.2e-17)
      const double xg = G_COULOMB * sqrt(dx * dx + dy * dy + dz * dz);
      uc += s[5] * G_COULOMB * (xg >= 6.0 ? 1.0 / xg : cl0(xg));
    }
    uc *= EPSILON0INV;
    double v[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < g.m; ++i) {
      const int64_t idx = i * g.P + e;
      const double ti = t[idx], bvi = bv[idx];
      const double c = closure_c<double>(g.closure, g.pse_order, bvi, ti);
      const double h = c + ti;
      const double xx = -bvi + ti;
      const double cl = -g.beta * g.q[i] * uc;
      const double w = g.rw[i];  // rho_i / rho (1: one density)
      v[0] += w * form_mu(g.closure, g.pse_order, xx, h, c, cl);
      v[1] += w * form_mu(0, 0, xx, h, c, cl);
      v[2] += w * form_mu(1, 1, xx, h, c, cl);
      v[3] += w * form_mu(2, 0, xx, h, c, cl);
      v[4] += w * c;
      if (go) go[idx] = h + 1.0;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {  // Neumaier
      const double s = sum[q] + v[q];
      comp[q] += (fabs(sum[q]) >= fabs(v[q])) ? (sum[q] - s) + v[q] : (v[q] - s) + sum[q];
      sum[q] = s;
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double s = warp_sum(sum[q] + comp[q]);
    if (lane == 0) red[q][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double s = red[threadIdx.x][0];
    for (int i = 1; i < NT / 64; ++i) s += red[threadIdx.x][i];
    part[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 5 + threadIdx.x] = s;
  }
}

// ------------------------------------------------------------ gradients
// Derivatives of the capped pair functions (bgy3d-solutes.c:102-131,
// 162-190, 246-305).
__device__ __forceinline__ double ljcap1(double r) {
  const double R = 0.2;
  if (r >= R) return lj1(r);
  return 2 * (lj1(R) / (2 * R)) * r;
}
__device__ __forceinline__ double cscap1(double r) {
  const double R = 0.2;
  if (r >= R) return cs1(r);
  return 2 * (cs1(R) / (2 * R)) * r;
}
__device__ __forceinline__ double cl1(double r) {
  const double r2 = r * r;
  if (r >= 0.02) return ((2 / sqrt(M_PI)) * r * exp(-r2) - erf(r)) / r2;
  return -(4 / (3 * sqrt(M_PI))) * r * (1. + r2 * (-3. / 5. + r2 * (3. / 14. + r2 * (-1. / 18.))));
}
__device__ __forceinline__ double ccap1(double r) {
  if (r >= 0.2) return -1.0 / (r * r);
  return cscap1(r) + cl1(r);
}

// d c / d t and d c / d(beta v) of closure_c, FP64.
__device__ __forceinline__ void closure_derivs(int method, int order, double bv, double t,
                                               double* ct, double* cv) {
  const double x = -bv + t;
  switch (method) {
    case 0: *ct = expm1(x); *cv = -(*ct + 1); return;
    case 1:
      if (x <= 0) { *ct = expm1(x); *cv = -(*ct + 1); } else { *ct = 0; *cv = -1; }
      return;
    case 2: *ct = exp(-bv) - 1; *cv = -exp(-bv) * (1 + t); return;
    default: {
      *ct = (x < 0) ? expm1(x) : exps1(order - 1, x);
      *cv = -(*ct + 1);
    }
  }
}

// form1 (closures.f90:615-653): differential of the mu density.
__device__ __forceinline__ double form1_mu(int method, int order, double x, double dx, double h,
                                           double dh, double c, double dc, double cl) {
  const double base = -dc - dh * (c + cl) / 2 - h * dc / 2;
  switch (method) {
    case 0: return base + h * dh;
    case 1: return (h < 0) ? base + h * dh : base;
    case 2: return base;
    default: {
      if (h < 0) return base + h * dh;
      double nx = 1, xn = 1;
      for (int k = 1; k <= order; ++k) { nx *= k; xn *= x; }
      return base + (h * dh - xn * dx / nx);
    }
  }
}

// Per-point coefficients at the converged t (grid-stride, one solute).
// mode 0 (TPT, hnc3d.c:1666-1742): A_i = rho h^3 g_i, B = 0.
// mode 1 (adjoint, prep): D_i = dc/dt, gmu_i = dmu/dt.
// mode 2 (adjoint, final): A_i = beta (K lambda)_i dc/d(beta v) + dmu/dv_i,
//                          B = rho h^3 / 2 sum_i q_i h_i  (dmu/du_c).
// mode 4 (linear response, hnc3d.c:1969-2134): as mode 2 with the
//   reference's dc/d(beta v) = -(dc/dt + 1) (its Jacobian flag hack,
//   hnc3d.c:1310-1320) for every closure; B unused.
__global__ void __launch_bounds__(NT)
adj_coef(GridArgs g, SitesDesc sd, const double* __restrict__ t, const double* __restrict__ bv,
         const double* __restrict__ Klam, double L0, double L1, double L2, double h0, double h1,
         double h2, double rho, int mode, double* __restrict__ A, double* __restrict__ B,
         double* __restrict__ D, double* __restrict__ gmu) {
  const double fmu = rho * g.h3 / g.beta;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.P; e += stride) {
    double uc = 0.0;
    if (mode != 0) {
      const int x = (int)(e % g.N0);
      const int y = (int)((e / g.N0) % g.N1);
      const int z = (int)(e / ((int64_t)g.N0 * g.N1));
      const double rx = x * h0 - L0 / 2, ry = y * h1 - L1 / 2, rz = z * h2 - L2 / 2;
      for (int a = 0; a < sd.n; ++a) {
        const double* s = sd.xyz_s_e_q + 6 * a;
        const double dx = rx - s[0], dy = ry - s[1], dz = rz - s[2];
        uc += s[5] * G_COULOMB * cl0(G_COULOMB * sqrt(dx * dx + dy * dy + dz * dz));
      }
      uc *= EPSILON0INV;
    }
    double bsum = 0.0;
    for (int i = 0; i < g.m; ++i) {
      const int64_t q = i * g.P + e;
      const double ti = t[q], bvi = bv[q];
      const double c = closure_c<double>(g.closure, g.pse_order, bvi, ti);
      const double h = c + ti;
      if (mode == 0) {
        A[q] = rho * g.h3 * (h + 1.0);
        continue;
      }
      double ct, cv;
      closure_derivs(g.closure, g.pse_order, bvi, ti, &ct, &cv);
      const double x = -bvi + ti;
      const double cl = -g.beta * g.q[i] * uc;
      if (mode == 1) {
        D[q] = ct;
        gmu[q] = fmu * form1_mu(g.closure, g.pse_order, x, 1.0, h, ct + 1, c, ct, cl);
      } else {
        const double dcv = (mode == 4 ? -(ct + 1) : cv) * g.beta;
        A[q] = Klam[q] * dcv +
               fmu * form1_mu(g.closure, g.pse_order, x, -g.beta, h, dcv, c, dcv, cl);
        bsum += g.q[i] * h;
      }
    }
    if (mode == 2) B[e] = 0.5 * rho * g.h3 * bsum;
  }
}

// sum_r sum_i A_i(r) dv_i(r)/dtheta_a (+ B(r) du_c/dtheta_a) for solute
// site a = blockIdx.y over a chunk of points blockIdx.x.  mode 0 (TPT,
// field1 bgy3d-solutes.c:381-468): full capped Coulomb, xyz only.
// mode 1 (adjoint): short-range v plus the u_c of the mu integral; x, y, z,
// q, sigma, epsilon.  Partials part[(a * gridDim.x + chunk) * 6 + p].
__global__ void __launch_bounds__(NT)
grad_field(GridArgs g, SitesDesc sd, const double* __restrict__ solvent_sse,
           const double* __restrict__ A, const double* __restrict__ B, double L0, double L1,
           double L2, double h0, double h1, double h2, int mode, double* __restrict__ part) {
  __shared__ double red[6][NT / 64];
  const int a = blockIdx.y;
  const double* s = sd.xyz_s_e_q + 6 * a;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * NT;
  const double G = G_COULOMB;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.P; e += stride) {
    const int x = (int)(e % g.N0);
    const int y = (int)((e / g.N0) % g.N1);
    const int z = (int)(e / ((int64_t)g.N0 * g.N1));
    const double dx = x * h0 - L0 / 2 - s[0], dy = y * h1 - L1 / 2 - s[1], dz = z * h2 - L2 / 2 - s[2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double inv = r > 0 ? 1.0 / r : 0.0;
    const double gx = -dx * inv, gy = -dy * inv, gz = -dz * inv;  // dr / dx_a
    double fr = 0.0;
    for (int i = 0; i < g.m; ++i) {
      const double* sv = solvent_sse + 3 * i;
      const double eab = sqrt(sv[1] * s[4]);
      const double sab = 0.5 * (sv[0] + s[3]);
      const double qab = sv[2] * s[5] * EPSILON0INV;
      const double Ai = A[i * g.P + e];
      const double rs = r / sab;
      const double lj1v = ljcap1(rs);
      if (mode == 0) {
        fr += Ai * ((eab / sab) * lj1v + qab * G * G * ccap1(G * r));
      } else {
        fr += Ai * ((eab / sab) * lj1v + qab * G * G * cscap1(G * r));
        acc[3] += Ai * sv[2] * EPSILON0INV * G * cscap0(G * r);
        acc[4] += Ai * eab * lj1v * (-r / (sab * sab)) * 0.5;
        if (eab > 0) acc[5] += Ai * ljcap0(rs) * sv[1] / (2 * eab);
      }
    }
    if (mode != 0) {
      const double b = B[e];
      fr += b * EPSILON0INV * s[5] * G * G * cl1(G * r);
      acc[3] += b * EPSILON0INV * G * cl0(G * r);
    }
    acc[0] += fr * gx;
    acc[1] += fr * gy;
    acc[2] += fr * gz;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    const double v = warp_sum(acc[p]);
    if (lane == 0) red[p][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = red[threadIdx.x][0];
    for (int i = 1; i < NT / 64; ++i) v += red[threadIdx.x][i];
    part[((int64_t)a * gridDim.x + blockIdx.x) * 6 + threadIdx.x] = v;
  }
}

// tau term of the adjoint gradient: with S(k) = w(kx) s(k) sum_i tau_i(k)
// conj(Lambda_i(k)) (Lambda = r2c(lambda), w = 1 at kx = 0 and N0/2, else
// 2), -tau_scale Re sum_k S(k) dF(k)/dtheta_a for F = sum_a Q_a e^{-ik.x_a}:
// theta = Q_a -> e^{-ik.x_a}, x_a -> -i k Q_a e^{-ik.x_a}.  Site a =
// blockIdx.y, chunk of K = blockIdx.x.  part[(a * gridDim.x + chunk) * 4 + p]
// for p = x, y, z, q.
__global__ void __launch_bounds__(NT)
grad_tau(GridArgs g, SitesDesc sd, const cpx<double>* __restrict__ Lam,
         const double* __restrict__ tau3, const double* __restrict__ chi3,
         double* __restrict__ part) {
  __shared__ double red[4][NT / 64];
  const int a = blockIdx.y;
  const double* s = sd.xyz_s_e_q + 6 * a;
  double acc[4] = {0, 0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.K; e += stride) {
    const int x = (int)(e % g.NH);
    if (x >= g.NX) continue;  // row padding
    const int y = (int)((e / g.NH) % g.N1);
    const int z = (int)(e / ((int64_t)g.NH * g.N1));
    const int iy = (y <= g.N1 / 2) ? y : y - g.N1;
    const int iz = (z <= g.N2 / 2) ? z : z - g.N2;
    const double kx = x * g.dkx, ky = iy * g.dky, kz = iz * g.dkz;
    // the long-range weight of site i: tau_i(k) on the tau path; on the Ng
    // path (whose residual carries -beta u_c(k) (q_i + sum_j (chi - 1)_ij q_j)
    // with u_c = eps0^-1 4 pi e^{-k^2/4G^2}/k^2 F s, 0 at k = 0) the same
    // factor without eps0^-1 F s, which tau_scale and the phase supply
    double shape = 0.0;
    if (!g.renorm) {
      const double k2 = kmag2(kx, ky, kz);
      shape = (k2 == 0.0) ? 0.0 : 4 * M_PI * exp(-k2 / (4 * G_COULOMB * G_COULOMB)) / k2;
    }
    double sre = 0.0, sim = 0.0;
    for (int i = 0; i < g.m; ++i) {
      const cpx<double> L = Lam[i * g.K + e];
      double ti;
      if (g.renorm) {
        ti = tau3[i * g.K + e];
      } else {
        double w = g.q[i];
        for (int j = 0; j < g.m; ++j)
          w += chi3[(int64_t)((i >= j) ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i) * g.K + e] *
               g.q[j];
        ti = shape * w;
      }
      sre += ti * L.re;
      sim -= ti * L.im;
    }
    const double w = ((x == 0 || x == g.N0 / 2) ? 1.0 : 2.0) * (((x + iy + iz) & 1) ? -1.0 : 1.0);
    sre *= w;
    sim *= w;
    double sn, cs;
    sincos(kx * s[0] + ky * s[1] + kz * s[2], &sn, &cs);
    // S e^{-i phi} = (sre + i sim)(cs - i sn)
    const double pre = sre * cs + sim * sn;
    const double pim = sim * cs - sre * sn;
    acc[3] += pre;                       // Re S e^{-i phi}
    // Re[S e^{-i phi} (-i k Q)] = Q k Im[S e^{-i phi}]
    acc[0] += s[5] * kx * pim;
    acc[1] += s[5] * ky * pim;
    acc[2] += s[5] * kz * pim;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const double v = warp_sum(acc[p]);
    if (lane == 0) red[p][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double v = red[threadIdx.x][0];
    for (int i = 1; i < NT / 64; ++i) v += red[threadIdx.x][i];
    part[((int64_t)a * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = -g.tau_scale * v;
  }
}

// ------------------------------------------------------ Newton-Krylov
// dc/dt at t (closure1, closures.f90:131-153) for the matrix-free Newton
// Jacobian, FP64, one solute.
__global__ void __launch_bounds__(NT)
closure_dcdt(GridArgs g, const double* __restrict__ t, const double* __restrict__ bv,
             double* __restrict__ D, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride) {
    double ct, cv;
    closure_derivs(g.closure, g.pse_order, bv[e], t[e], &ct, &cv);
    D[e] = ct;
  }
}

// Krylov helpers, FP64, one solute.  mdot: per workgroup partial sums of
// <w, V_i>, V_i = V + i n for i < nv, to part[block * nv + i] (the host adds
// the partials in a fixed order).  maxpy: w -= sum_i h_i V_i.  vcomb:
// y = a x + b y.
constexpr int KMAX = 32;
__global__ void __launch_bounds__(NT)
mdot(const double* __restrict__ V, int nv, const double* __restrict__ w, int64_t n,
     double* __restrict__ part) {
  __shared__ double red[KMAX][NT / 64];
  double acc[KMAX];
#pragma unroll
  for (int i = 0; i < KMAX; ++i) acc[i] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride) {
    const double we = w[e];
#pragma unroll
    for (int i = 0; i < KMAX; ++i)
      if (i < nv) acc[i] += we * V[(int64_t)i * n + e];
  }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < KMAX; ++i)
    if (i < nv) {
      const double v = warp_sum(acc[i]);
      if (lane == 0) red[i][wv] = v;
    }
  __syncthreads();
  if (threadIdx.x < nv) {
    double v = red[threadIdx.x][0];
    for (int q = 1; q < NT / 64; ++q) v += red[threadIdx.x][q];
    part[(int64_t)blockIdx.x * nv + threadIdx.x] = v;
  }
}
__global__ void __launch_bounds__(NT)
maxpy(double* __restrict__ w, const double* __restrict__ V, const double* __restrict__ h, int nv,
      int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride) {
    double acc = w[e];
    for (int i = 0; i < nv; ++i) acc -= h[i] * V[(int64_t)i * n + e];
    w[e] = acc;
  }
}
__global__ void __launch_bounds__(NT)
vcomb(double* __restrict__ y, const double* __restrict__ x, double a, double b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride)
    y[e] = a * x[e] + (b == 0.0 ? 0.0 : b * y[e]);
}

// ------------------------------------------------ 3D solvent-solvent
// beta v_short of every site pair around the grid centre
// (bgy3d_pair_potential, bgy3d-force.c:123-143): LJ with its epsilon *
// 1e8 ceiling and erfc Coulomb with its 1e3 ceiling (bgy3d-force.h:23-88),
// r^2 summed x, y, z (vec_rmap, bgy3d-vec.h:639-648).  pse: [msol][3]
// sigma, epsilon, charge.  Grid: x = chunks of P, y = pair.
__global__ void __launch_bounds__(NT)
pair_field(GridArgs g, const double* __restrict__ sse, double L0, double L1, double L2,
           double h0, double h1, double h2, double* __restrict__ bv) {
  const int p = blockIdx.y;
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= p) ++i;
  const int j = p - i * (i + 1) / 2;
  const double eps = sqrt(sse[3 * i + 1] * sse[3 * j + 1]);
  const double sig = 0.5 * (sse[3 * i] + sse[3 * j]);
  const double q2 = sse[3 * i + 2] * sse[3 * j + 2];
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.P; e += stride) {
    const int x = (int)(e % g.N0);
    const int y = (int)((e / g.N0) % g.N1);
    const int z = (int)(e / ((int64_t)g.N0 * g.N1));
    const double rx = x * h0 - L0 / 2, ry = y * h1 - L1 / 2, rz = z * h2 - L2 / 2;
    const double r = sqrt((rx * rx + ry * ry) + rz * rz);
    double lj;
    if ((sig == 0.0 || eps == 0.0) && r == 0.0) {
      lj = 0.0;
    } else {
      const double sr = sig / r;
      const double sr6 = (sr * sr) * (sr * sr) * (sr * sr);
      const double re = 4 * eps * sr6 * (sr6 - 1);
      lj = (fabs(re) > eps * 1.0e8) ? eps * 1.0e8 : re;
    }
    const double f = erfc(G_COULOMB * r) / r;
    const double cs = (f > 1.0e3) ? EPSILON0INV * q2 * 1.0e3 : EPSILON0INV * q2 * f;
    bv[p * g.P + e] = g.beta * (lj + cs);
  }
}

// Intramolecular omega_ij(k) = sinc(2 pi r_ij / L0 sqrt(k2)), k2 the
// integer kz^2 + ky^2 + kx^2 (omega_intra, bgy3d-pure.c:236-267), into
// w[p * K + e] for the off-diagonal pairs.  rij: [npairs].  Grid: x =
// chunks of K, y = pair.
__global__ void __launch_bounds__(NT)
omega3(GridArgs g, double L0, const double* __restrict__ rij, double* __restrict__ w) {
  const int p = blockIdx.y;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.K; e += stride) {
    const int x = (int)(e % g.NH);
    if (x >= g.NX) continue;
    const int y = (int)((e / g.NH) % g.N1);
    const int z = (int)(e / ((int64_t)g.NH * g.N1));
    const int iy = (y <= g.N1 / 2) ? y : y - g.N1;
    const int iz = (z <= g.N2 / 2) ? z : z - g.N2;
    const int k2 = iz * iz + iy * iy + x * x;
    const double kr = (2.0 * M_PI * rij[p] / L0) * sqrt((double)k2);
    w[p * g.K + e] = (kr == 0.0) ? 1.0 : sin(kr) / kr;
  }
}

// mu densities of the solvent-solvent pairs (print_chempot m x m,
// hnc3d.c:486-623, 669-687), off-diagonal pairs counted twice: native,
// HNC, KH, GF and, with gout, g = h + 1 per pair.  uL: the unit long-range
// Coulomb in real space, cl = -beta q_i q_j uL.  Neumaier sums per thread,
// part[block * 4 + q].  Grid: x = chunks of P.
__global__ void __launch_bounds__(NT)
mu_pairs(GridArgs g, const double* __restrict__ t, const double* __restrict__ bv,
         const double* __restrict__ uL, double* __restrict__ gout, double* __restrict__ part) {
  __shared__ double red[4][NT / 64];
  double sum[4] = {0, 0, 0, 0}, comp[4] = {0, 0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.P; e += stride) {
    double v[4] = {0, 0, 0, 0};
    for (int i = 0; i < g.msol; ++i)
      for (int j = 0; j <= i; ++j) {
        const int p = i * (i + 1) / 2 + j;
        const int64_t idx = p * g.P + e;
        const double ti = t[idx], bvi = bv[idx];
        const double c = closure_c<double>(g.closure, g.pse_order, bvi, ti);
        const double h = c + ti;
        const double xx = -bvi + ti;
        const double cl = -g.beta * g.q[i] * g.q[j] * uL[e];
        const double mult = (i == j) ? 1.0 : 2.0;
        v[0] += mult * form_mu(g.closure, g.pse_order, xx, h, c, cl);
        v[1] += mult * form_mu(0, 0, xx, h, c, cl);
        v[2] += mult * form_mu(1, 1, xx, h, c, cl);
        v[3] += mult * form_mu(2, 0, xx, h, c, cl);
        if (gout) gout[idx] = h + 1.0;
      }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double s = sum[q] + v[q];
      comp[q] += (fabs(sum[q]) >= fabs(v[q])) ? (sum[q] - s) + v[q] : (v[q] - s) + sum[q];
      sum[q] = s;
    }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double s = warp_sum(sum[q] + comp[q]);
    if (lane == 0) red[q][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = red[threadIdx.x][0];
    for (int i = 1; i < NT / 64; ++i) s += red[threadIdx.x][i];
    part[(int64_t)blockIdx.x * 4 + threadIdx.x] = s;
  }
}

// chi - 1 of the solvent at the corner from the spectrum W_p = FFT(rho h_p)
// (hnc3d.c:1006-1052): h^3 W s + (omega_p - 1) with omega_ii = 1; out is
// complex [npairs][N2][N1][NX] without the row padding.  Grid: x = chunks
// of K, y = pair.
__global__ void __launch_bounds__(NT)
chi_from_h(GridArgs g, const cpx<double>* __restrict__ W, const double* __restrict__ omega,
           double* __restrict__ out) {
  const int p = blockIdx.y;
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= p) ++i;
  const int j = p - i * (i + 1) / 2;
  const int64_t stride = (int64_t)gridDim.x * NT;
  const int64_t Kv = (int64_t)g.NX * g.N1 * g.N2;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.K; e += stride) {
    const int x = (int)(e % g.NH);
    if (x >= g.NX) continue;
    const int y = (int)((e / g.NH) % g.N1);
    const int z = (int)(e / ((int64_t)g.NH * g.N1));
    const int iy = (y <= g.N1 / 2) ? y : y - g.N1;
    const int iz = (z <= g.N2 / 2) ? z : z - g.N2;
    const double sgn = ((x + iy + iz) & 1) ? -1.0 : 1.0;
    const cpx<double> wv = W[p * g.K + e];
    double re = wv.re * g.h3 * sgn, im = wv.im * g.h3 * sgn;
    if (i != j) re += omega[p * g.K + e];
    const int64_t o = p * Kv + ((int64_t)z * g.N1 + y) * g.NX + x;
    out[2 * o] = re;
    out[2 * o + 1] = im;
  }
}

// FP64 -> FP32 copy.
__global__ void __launch_bounds__(NT)
to_f32(float* __restrict__ dst, const double* __restrict__ src, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < n; e += stride)
    dst[e] = (float)src[e];
}

// ------------------------------------------------------ info observables
// Solvent charge density rho_v = sum_i (q_i rho) g_i with g_i = (c_i + t_i)
// + 1 (bgy3d_solvent_field, bgy3d-potential.c:246-262, on the g that
// hnc3d.c:2319-2325 passes).  One solute; grid-stride over P.
__global__ void __launch_bounds__(NT)
info_rho_v(GridArgs g, const double* __restrict__ t, const double* __restrict__ bv, double rho,
           double* __restrict__ rhov) {
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.P; e += stride) {
    double acc = 0.0;
    for (int i = 0; i < g.m; ++i) {
      const int64_t idx = i * g.P + e;
      const double c = closure_c<double>(g.closure, g.pse_order, bv[idx], t[idx]);
      acc += (g.q[i] * rho * g.rw[i]) * ((c + t[idx]) + 1.0);
    }
    rhov[e] = acc;
  }
}

// The grid sums of print_info (bgy3d-potential.c:341-443), unscaled, with
// X_a = i_a - N_a / 2 (bgy3d_moments, bgy3d-vec.c:573-593):
//   [0..9]   rho_v: 1, X, Y, Z, XX, XY, XZ, YY, YZ, ZZ
//   [10..19] rho_u, the Gaussian cores (G/sqrt(pi))^3 sum_a q_a
//            exp(-G^2 r_a^2) (bgy3d-solutes.c:555-577), same moments
//   [20] U_v rho_u   [21] U_u rho_v, U_u = eps0^-1 sum_a q_a G cl0(G r_a)
// Neumaier sums per thread, part[block * INFO_NSUM + q].  Grid: x = chunks
// of P.
__global__ void __launch_bounds__(NT)
info_sums(GridArgs g, SitesDesc sd, const double* __restrict__ rhov,
          const double* __restrict__ ve, double L0, double L1, double L2, double h0, double h1,
          double h2, double* __restrict__ part) {
  __shared__ double red[INFO_NSUM][NT / 64];
  double sum[INFO_NSUM], comp[INFO_NSUM];
#pragma unroll
  for (int q = 0; q < INFO_NSUM; ++q) { sum[q] = 0.0; comp[q] = 0.0; }
  const double gn = G_COULOMB / sqrt(M_PI);
  const double norm = gn * gn * gn;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.P; e += stride) {
    const int x = (int)(e % g.N0);
    const int y = (int)((e / g.N0) % g.N1);
    const int z = (int)(e / ((int64_t)g.N0 * g.N1));
    const double rx = x * h0 - L0 / 2, ry = y * h1 - L1 / 2, rz = z * h2 - L2 / 2;
    double ru = 0.0, uc = 0.0;
    for (int a = 0; a < sd.n; ++a) {
      const double* s = sd.xyz_s_e_q + 6 * a;
      const double dx = rx - s[0], dy = ry - s[1], dz = rz - s[2];
      const double r2 = (dx * dx + dy * dy) + dz * dz;
      ru += s[5] * exp(-G_COULOMB * G_COULOMB * r2);
      uc += s[5] * G_COULOMB * cl0(G_COULOMB * sqrt(r2));
    }
    ru *= norm;
    uc *= EPSILON0INV;
    const double X = x - 0.5 * g.N0, Y = y - 0.5 * g.N1, Z = z - 0.5 * g.N2;
    const double rv = rhov[e];
    const double w[10] = {1.0, X, Y, Z, X * X, X * Y, X * Z, Y * Y, Y * Z, Z * Z};
    double v[INFO_NSUM];
#pragma unroll
    for (int q = 0; q < 10; ++q) { v[q] = rv * w[q]; v[10 + q] = ru * w[q]; }
    v[20] = ve[e] * ru;
    v[21] = uc * rv;
#pragma unroll
    for (int q = 0; q < INFO_NSUM; ++q) {
      const double s = sum[q] + v[q];
      comp[q] += (fabs(sum[q]) >= fabs(v[q])) ? (sum[q] - s) + v[q] : (v[q] - s) + sum[q];
      sum[q] = s;
    }
  }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < INFO_NSUM; ++q) {
    const double s = warp_sum(sum[q] + comp[q]);
    if (lane == 0) red[q][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x < INFO_NSUM) {
    double s = red[threadIdx.x][0];
    for (int i = 1; i < NT / 64; ++i) s += red[threadIdx.x][i];
    part[(int64_t)blockIdx.x * INFO_NSUM + threadIdx.x] = s;
  }
}

// The solvent potential at the solute sites by trigonometric interpolation
// (bgy3d_pot_interp, bgy3d-potential.c:125-143, and bgy3d_fft_interp,
// bgy3d-fftw3.c:504-612) straight from Y = FFT(rho_v) (spectrum of site 0
// in W, padded rows): FFT(U_v) = N0 N1 N2 Y f / L^3 with the Poisson factor
// f = 4 pi eps0^-1 h^3 / k^2 (0 at k = 0), so
//   U_v(y) = sum_k 2 Re(Y f / L^3 w_x w_y w_z) / (2 if k_x = 0 or N0/2),
// w_d = exp(2 pi i K_d y_d / N_d), cos(...) where K_d = -K_d mod N_d, and
// y_d = (x_d + L_d / 2) / h_d.  xs: [n][3].  part[(a * gridDim.x + block)].
// Grid: x = chunks of K, y = site.
__global__ void __launch_bounds__(NT)
info_interp(GridArgs g, const cpx<double>* __restrict__ W, const double* __restrict__ xs,
            double L0, double L1, double L2, double h0, double h1, double h2,
            double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int a = blockIdx.y;
  const double yx = (xs[3 * a] + L0 / 2) / h0;
  const double yy = (xs[3 * a + 1] + L1 / 2) / h1;
  const double yz = (xs[3 * a + 2] + L2 / 2) / h2;
  double sum = 0.0, comp = 0.0;
  const int64_t stride = (int64_t)gridDim.x * NT;
  for (int64_t e = blockIdx.x * (int64_t)NT + threadIdx.x; e < g.K; e += stride) {
    const int x = (int)(e % g.NH);
    if (x >= g.NX) continue;
    const int y = (int)((e / g.NH) % g.N1);
    const int z = (int)(e / ((int64_t)g.NH * g.N1));
    const int iy = (y <= g.N1 / 2) ? y : y - g.N1;
    const int iz = (z <= g.N2 / 2) ? z : z - g.N2;
    const double kx = x * g.dkx, ky = iy * g.dky, kz = iz * g.dkz;
    const double k2 = (kz * kz + ky * ky) + kx * kx;
    if (k2 == 0.0) continue;
    const double f = 4 * M_PI * EPSILON0INV * g.h3 / k2 * g.inv_L3;
    // the three one-dimensional waves
    double cx, sx, cy, sy, cz, sz;
    sincos((2 * M_PI / g.N0) * x * yx, &sx, &cx);
    sincos((2 * M_PI / g.N1) * iy * yy, &sy, &cy);
    sincos((2 * M_PI / g.N2) * iz * yz, &sz, &cz);
    if ((g.N0 - 2 * x) % g.N0 == 0) sx = 0.0;
    if ((g.N1 - 2 * y) % g.N1 == 0) sy = 0.0;
    if ((g.N2 - 2 * z) % g.N2 == 0) sz = 0.0;
    // w = wx wy wz
    const double ar = cx * cy - sx * sy, ai = cx * sy + sx * cy;
    const double wr = ar * cz - ai * sz, wi = ar * sz + ai * cz;
    const cpx<double> Yk = W[e];
    double v = 2.0 * f * (Yk.re * wr - Yk.im * wi);
    if ((g.N0 - 2 * x) % g.N0 == 0) v *= 0.5;
    const double s = sum + v;
    comp += (fabs(sum) >= fabs(v)) ? (sum - s) + v : (v - s) + sum;
    sum = s;
  }
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double s = warp_sum(sum + comp);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = red[0];
    for (int i = 1; i < NT / 64; ++i) acc += red[i];
    part[(int64_t)a * gridDim.x + blockIdx.x] = acc;
  }
}

}  // namespace rism
