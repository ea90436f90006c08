// Host side of the MI355X 3D-RISM solver: workspace, kernel dispatch,
// the batched MDIIS / Picard loop, and the C ABI of include/rism3d.h.
//
// Reference counterparts: hnc3d_solute_solve (hnc3d.c:1762-2350),
// solvent_kernel_rism (hnc3d.c:1443-1566), rism_solvent_renorm
// (rism.f90:141-208), chempot / pmv_factor (hnc3d.c:604-737),
// bgy3d_snes_picard (bgy3d-snes.c:348-384).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rism3d.h"

#include "kernels.hip"

// The fixed device scratch of rism3d_residual_noalloc: every device buffer
// of that entry point is carved from this array, of which each device holds
// its own copy from module load to process end.  Host code reaches it only
// through the address hipGetSymbolAddress returns on each device; the name
// itself is not a device address on the host.
static __device__ __align__(256) unsigned char scratch[(size_t)RISM3D_NOALLOC_SCRATCH_MB << 20];

namespace rism {
namespace {

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "rism3d: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                                  \
      throw HipFailure{e_};                                                              \
    }                                                                                    \
  } while (0)

struct HipFailure {
  hipError_t err;
};

bool supported_n(int n) {
  switch (n) {
    case 32: case 48: case 64: case 96: case 128: case 192: case 256: case 384: return true;
    default: return false;
  }
}

// ----------------------------------------------------- kernel dispatch
#define RISM_FOR_N(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384)

// The history loads of the closure and residual kernels are unrolled over
// H slots held in registers: the smallest of 8, 10 and MAX_DIIS that holds
// every solute's history (at 10 slots the residual kernel keeps 80 fewer
// bytes per lane in flight than at 16).
template <typename T>
void launch_closure_r2c(int N0, dim3 grid, hipStream_t s, const SoluteDesc* sd, const GridArgs& g,
                        const cpx<T>* twh, const cpx<T>* twf, int hist) {
  if (hist <= 8) {
    switch (N0) {
#define X(n) case n: hipLaunchKernelGGL((closure_r2c_x<T, n, 8>), grid, dim3(NT), 0, s, sd, g, twh, twf); break;
      RISM_FOR_N(X)
#undef X
    }
    return;
  }
  if (hist <= 10) {
    switch (N0) {
#define X(n) case n: hipLaunchKernelGGL((closure_r2c_x<T, n, 10>), grid, dim3(NT), 0, s, sd, g, twh, twf); break;
      RISM_FOR_N(X)
#undef X
    }
    return;
  }
  switch (N0) {
#define X(n) case n: hipLaunchKernelGGL((closure_r2c_x<T, n, MAX_DIIS>), grid, dim3(NT), 0, s, sd, g, twh, twf); break;
    RISM_FOR_N(X)
#undef X
  }
}
template <typename T>
void launch_c2r_resid(int N0, dim3 grid, hipStream_t s, const SoluteDesc* sd, const GridArgs& g,
                      const cpx<T>* twh, const cpx<T>* twf, int hist) {
  if (hist <= 8) {
    switch (N0) {
#define X(n) case n: hipLaunchKernelGGL((c2r_x_resid<T, n, 8>), grid, dim3(NT), 0, s, sd, g, twh, twf); break;
      RISM_FOR_N(X)
#undef X
    }
    return;
  }
  if (hist <= 10) {
    switch (N0) {
#define X(n) case n: hipLaunchKernelGGL((c2r_x_resid<T, n, 10>), grid, dim3(NT), 0, s, sd, g, twh, twf); break;
      RISM_FOR_N(X)
#undef X
    }
    return;
  }
  switch (N0) {
#define X(n) case n: hipLaunchKernelGGL((c2r_x_resid<T, n, MAX_DIIS>), grid, dim3(NT), 0, s, sd, g, twh, twf); break;
    RISM_FOR_N(X)
#undef X
  }
}
template <typename T> constexpr int y_lines() { return 128 / (int)sizeof(cpx<T>); }
template <typename T, bool INV>
void launch_fft_y(int N1, dim3 grid, hipStream_t s, const SoluteDesc* sd, const GridArgs& g,
                  const cpx<T>* tw) {
  constexpr int LN = y_lines<T>();
  switch (N1) {
#define X(n) case n: hipLaunchKernelGGL((fft_y<T, n, LN, INV>), grid, dim3(NT), 0, s, sd, g, tw); break;
    RISM_FOR_N(X)
#undef X
  }
}
template <typename T, int LINES>
void launch_zstar_l(int N2, dim3 grid, size_t lds, hipStream_t s, const SoluteDesc* sd,
                    const GridArgs& g, const cpx<T>* tw, const T* chi, const T* tau) {
  switch (N2) {
#define X(n) case n: hipLaunchKernelGGL((zstar<T, n, LINES>), grid, dim3(NT), lds, s, sd, g, tw, chi, tau); break;
    RISM_FOR_N(X)
#undef X
  }
}
// Tuning knobs, read once: RISM3D_ZSTAR_LINES (kx columns per zstar
// workgroup, 2..16) and RISM3D_ZSTAR_ORDER (0..3, see zstar_reg).
int env_int(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return v ? std::atoi(v) : dflt;
}
template <typename T>
int zstar_lines(int N2, int m) {
  // 8 kx columns per workgroup (64 B FP32, 128 B FP64 rows): at 128^3 the
  // FP32 pass ran 1.7 ms with 8 against 2.25 ms with 16 and 2.3 ms with 4
  // (64 solutes, m = 3; occupancy against row length)
  int L = 8;
  const int want = env_int("RISM3D_ZSTAR_LINES", 0);
  if (want == 2 || want == 4 || want == 8 || want == 16) L = want;
  while (L > 2 && (size_t)m * L * lstride_for(N2) * sizeof(cpx<T>) > 64 * 1024) L /= 2;
  return L;
}
template <typename T, int N2, int NTZ = NT>
void launch_zreg_n(int NX, int N1, int B, hipStream_t s, const SoluteDesc* sd, const GridArgs& g,
                   const cpx<T>* tw, const T* chi, const T* tau) {
  // NTZ threads: the same points per thread over NTZ / NT times the columns
  constexpr int L = zreg_lines<T, N2>() * (NTZ / NT);
  // In FP32 the Nyquist column kx = N0/2 gets a launch of its own (NYQ,
  // LINES rows per workgroup) when the tiles of LINES columns end just
  // before it, instead of a tile that holds it alone (one workgroup in five
  // at N0 = 128): 1.19 -> 1.08 ms per pass at 64 x 128^3 (2, 4, 8 or 16
  // rows per Nyquist workgroup alike).  In FP64, and in the y passes, the
  // separate launch cost more than the tile it replaced.  RISM3D_NYQ=0
  // keeps it in the last tile (g.nyq).
  const bool split = g.nyq && sizeof(T) == 4 && NX > 1 && (NX - 1) % L == 0;
  if (g.ylay && !split) {  // evaluate forces the split for the plane chain
    std::fprintf(stderr, "rism3d: y-major spectra need the Nyquist launch\n");
    throw HipFailure{hipErrorInvalidValue};
  }
  const int tiles = split ? (NX - 1) / L : (NX + L - 1) / L;
  GridArgs gz = g;
  if (gz.zorder == 2 && (tiles * N1) % 8 != 0) gz.zorder = 1;  // whole tiles per XCD
  if (gz.zorder == 3 && N1 % 8 != 0) gz.zorder = 1;             // whole rows per XCD
  gz.zB = B;
  gz.ztiles = tiles;
  const dim3 grid = gz.zorder >= 2 ? dim3(tiles * N1 * B)
                  : gz.zorder ? dim3(tiles, N1, B) : dim3(B, tiles, N1);
  const dim3 gn((N1 + L - 1) / L, 1, B);
  if (g.m == 3) {
    hipLaunchKernelGGL((zstar_reg<T, N2, L, 3, false, NTZ>), grid, dim3(NTZ), 0, s, sd, gz, tw, chi, tau);
    if (split) hipLaunchKernelGGL((zstar_reg<T, N2, L, 3, true, NTZ>), gn, dim3(NTZ), 0, s, sd, gz, tw, chi, tau);
  } else {
    hipLaunchKernelGGL((zstar_reg<T, N2, L, 5, false, NTZ>), grid, dim3(NTZ), 0, s, sd, gz, tw, chi, tau);
    if (split) hipLaunchKernelGGL((zstar_reg<T, N2, L, 5, true, NTZ>), gn, dim3(NTZ), 0, s, sd, gz, tw, chi, tau);
  }
}
// The FP32 residual on xy planes (closure_r2c_xy, zstar, c2r_xy_resid:
// three launches, the y transforms inside the plane kernels) when the
// planes are square with 64, 96 or 128 points per side; otherwise, and in
// FP64 (a 128^2 plane would take 133 KB of LDS), the five-launch chain.
// RISM3D_PLANES=0 keeps the five-launch chain in FP32 too.  The plane
// kernels write W y-major with the Nyquist column apart (GridArgs::ylay),
// which only zstar_reg reads (3 or 5 solvent sites).
bool use_zreg(const GridArgs& g);
bool use_planes(const GridArgs& g, bool fp32, int on) {
  return on && fp32 && g.special == 0 && g.N0 == g.N1 &&
         (g.N0 == 64 || g.N0 == 96 || g.N0 == 128) && use_zreg(g);
}
template <typename T, int N, int H>
void launch_planes_h(bool inv, dim3 grid, hipStream_t s, const SoluteDesc* sd, const GridArgs& g,
                     const cpx<T>* twh, const cpx<T>* twf, const cpx<T>* twy) {
  if (inv)
    hipLaunchKernelGGL((c2r_xy_resid<T, N, N, H>), grid, dim3(NTP), 0, s, sd, g, twh, twf, twy);
  else
    hipLaunchKernelGGL((closure_r2c_xy<T, N, N, H>), grid, dim3(NTP), 0, s, sd, g, twh, twf, twy);
}
template <typename T, int N>
void launch_planes_n(bool inv, dim3 grid, hipStream_t s, const SoluteDesc* sd, const GridArgs& g,
                     const cpx<T>* twh, const cpx<T>* twf, const cpx<T>* twy, int hist) {
  if (hist <= 8) launch_planes_h<T, N, 8>(inv, grid, s, sd, g, twh, twf, twy);
  else if (hist <= 10) launch_planes_h<T, N, 10>(inv, grid, s, sd, g, twh, twf, twy);
  else launch_planes_h<T, N, MAX_DIIS>(inv, grid, s, sd, g, twh, twf, twy);
}
// Grid: x = z plane, y = site, z = solute.
template <typename T>
void launch_planes(bool inv, int B, hipStream_t s, const SoluteDesc* sd, const GridArgs& g,
                   const cpx<T>* twh, const cpx<T>* twf, const cpx<T>* twy, int hist) {
  const dim3 grid(g.N2, g.m, B);
  switch (g.N0) {
    case 64: launch_planes_n<T, 64>(inv, grid, s, sd, g, twh, twf, twy, hist); break;
    case 96: launch_planes_n<T, 96>(inv, grid, s, sd, g, twh, twf, twy, hist); break;
    case 128: launch_planes_n<T, 128>(inv, grid, s, sd, g, twh, twf, twy, hist); break;
  }
}

// zstar_reg for the solute residual with 3 or 5 solvent sites (RISM3D_ZSTAR_LDS=1 keeps
// the all-sites-in-LDS kernel), zstar otherwise.
bool use_zreg(const GridArgs& g) {
  static const int lds_only = env_int("RISM3D_ZSTAR_LDS", 0);
  return !lds_only && g.special == 0 && (g.m == 3 || g.m == 5);
}
template <typename T>
void launch_zstar(int N2, int NX, int N1, int B, hipStream_t s, const SoluteDesc* sd,
                  const GridArgs& g, const cpx<T>* tw, const T* chi, const T* tau) {
  if (use_zreg(g)) {
    if (g.zwide && sizeof(T) == 4 && N2 <= 128) {  // 512 threads over 32 kx columns
      switch (N2) {
        case 32: launch_zreg_n<T, 32, 2 * NT>(NX, N1, B, s, sd, g, tw, chi, tau); break;
        case 48: launch_zreg_n<T, 48, 2 * NT>(NX, N1, B, s, sd, g, tw, chi, tau); break;
        case 64: launch_zreg_n<T, 64, 2 * NT>(NX, N1, B, s, sd, g, tw, chi, tau); break;
        case 96: launch_zreg_n<T, 96, 2 * NT>(NX, N1, B, s, sd, g, tw, chi, tau); break;
        case 128: launch_zreg_n<T, 128, 2 * NT>(NX, N1, B, s, sd, g, tw, chi, tau); break;
      }
      return;
    }
    switch (N2) {
#define X(n) case n: launch_zreg_n<T, n>(NX, N1, B, s, sd, g, tw, chi, tau); break;
      RISM_FOR_N(X)
#undef X
    }
    return;
  }
  const int L = zstar_lines<T>(N2, g.m);
  const size_t lds = (size_t)g.m * L * lstride_for(N2) * sizeof(cpx<T>);
  const dim3 grid = g.zorder ? dim3((NX + L - 1) / L, N1, B) : dim3(B, (NX + L - 1) / L, N1);
  switch (L) {
    case 2: launch_zstar_l<T, 2>(N2, grid, lds, s, sd, g, tw, chi, tau); break;
    case 4: launch_zstar_l<T, 4>(N2, grid, lds, s, sd, g, tw, chi, tau); break;
    case 8: launch_zstar_l<T, 8>(N2, grid, lds, s, sd, g, tw, chi, tau); break;
    default: launch_zstar_l<T, 16>(N2, grid, lds, s, sd, g, tw, chi, tau); break;
  }
}

// ------------------------------------------------------ device buffers
// Device memory is cached across solves: a buffer released by one solve is
// kept, keyed by device and byte size, and handed to the next request of
// that size, so a batch of the same shape allocates nothing after the
// first (hipMalloc / hipFree cost milliseconds per large buffer and
// synchronise the device).  Every solve runs on one stream and synchronises
// it before returning, so a released buffer is idle when handed out again.
// By default the cache keeps only what the last call used: at the end of
// each call the buffers it did not take are freed, so a change of shape or
// precision returns the memory of the old one.  RISM3D_CACHE_MB instead
// keeps every released buffer up to that many MB (0 disables the cache);
// rism3d_cache_release() empties it.  On an allocation failure the cache of
// that device is emptied and the allocation retried.
struct MemCache {
  std::mutex mu;
  // (device, bytes) -> (buffer, call that released it)
  std::multimap<std::pair<int, size_t>, std::pair<void*, uint64_t>> blocks;
  size_t bytes = 0;
  uint64_t call = 0;  // calls so far
  bool keep_last_only = std::getenv("RISM3D_CACHE_MB") == nullptr;
  size_t limit = keep_last_only ? SIZE_MAX
                                : (size_t)std::max(0, env_int("RISM3D_CACHE_MB", 0)) << 20;
};
MemCache& mem_cache() {
  static MemCache c;
  return c;
}
void cache_empty(int dev) {  // dev < 0: every device
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto it = c.blocks.begin(); it != c.blocks.end();) {
    if (dev >= 0 && it->first.first != dev) { ++it; continue; }
    (void)hipSetDevice(it->first.first);
    (void)hipFree(it->second.first);
    c.bytes -= it->first.second;
    it = c.blocks.erase(it);
  }
  (void)hipSetDevice(cur);
}
// End of a call: unless RISM3D_CACHE_MB asks to keep more, free the cached
// buffers this call did not release (those of earlier calls' shapes).
void cache_end_call() {
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.keep_last_only) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto it = c.blocks.begin(); it != c.blocks.end();) {
      if (it->second.second == c.call) { ++it; continue; }
      (void)hipSetDevice(it->first.first);
      (void)hipFree(it->second.first);
      c.bytes -= it->first.second;
      it = c.blocks.erase(it);
    }
    (void)hipSetDevice(cur);
  }
  ++c.call;
}
// A bump arena over memory the library already holds (the fixed scratch).
// While a thread has one installed, dev_alloc carves its buffers from it and
// dev_free leaves them there: the arena is released as a whole when the
// call that installed it returns.  A request beyond its end throws
// ScratchFull before anything is written.
struct ScratchFull {};
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
};
thread_local Arena* t_arena = nullptr;

void* dev_alloc(size_t bytes) {
  if (Arena* a = t_arena) {
    const size_t off = (a->used + 255) & ~(size_t)255;
    if (off > a->size || bytes > a->size - off) throw ScratchFull{};
    a->used = off + bytes;
    return a->base + off;
  }
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  MemCache& c = mem_cache();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.blocks.find({dev, bytes});
    if (it != c.blocks.end()) {
      void* p = it->second.first;
      c.blocks.erase(it);
      c.bytes -= bytes;
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    cache_empty(dev);
    HIPCHK(hipMalloc(&p, bytes));
  }
  return p;
}
void dev_free(void* p, size_t bytes) {
  if (const Arena* a = t_arena)
    if ((char*)p >= a->base && (char*)p < a->base + a->size) return;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  MemCache& c = mem_cache();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.bytes + bytes <= c.limit) {
      c.blocks.insert({{dev, bytes}, {p, c.call}});
      c.bytes += bytes;
      return;
    }
  }
  (void)hipFree(p);
}

template <typename X>
struct DevBuf {
  X* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) : n(count) {
    if (count) p = static_cast<X*>(dev_alloc(count * sizeof(X)));
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) dev_free(p, n * sizeof(X));
    p = nullptr;
    n = 0;
  }
};

// Pinned host staging for the small per-iteration transfers (solute
// descriptors up, residual norms and Gram rows down): pageable copies are
// staged by the runtime and block the host.  A ring of slots, each reused
// only after the copy that last read it has completed.
struct PinnedRing {
  static constexpr int SLOTS = 4;
  void* host[SLOTS] = {};
  hipEvent_t ev[SLOTS] = {};
  bool used[SLOTS] = {};
  size_t bytes = 0;
  int next = 0;
  ~PinnedRing() {
    for (int k = 0; k < SLOTS; ++k) {
      if (host[k]) (void)hipHostFree(host[k]);
      if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
  }
  void* slot(size_t need, int* k_out) {
    if (need > bytes) {
      for (int k = 0; k < SLOTS; ++k) {
        if (used[k]) HIPCHK(hipEventSynchronize(ev[k]));
        if (host[k]) HIPCHK(hipHostFree(host[k]));
        host[k] = nullptr;
        used[k] = false;
      }
      for (int k = 0; k < SLOTS; ++k) {
        HIPCHK(hipHostMalloc(&host[k], need, hipHostMallocDefault));
        if (!ev[k]) HIPCHK(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
      }
      bytes = need;
    }
    const int k = next;
    next = (next + 1) % SLOTS;
    if (used[k]) HIPCHK(hipEventSynchronize(ev[k]));
    *k_out = k;
    return host[k];
  }
  void mark(int k, hipStream_t s) {
    HIPCHK(hipEventRecord(ev[k], s));
    used[k] = true;
  }
};

template <typename T>
std::vector<cpx<T>> twiddles(int n) {
  std::vector<cpx<T>> w(n);
  for (int j = 0; j < n; ++j) {
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * j / n;
    w[j] = cpx<T>{(T)std::cos(a), (T)std::sin(a)};
  }
  return w;
}

struct Twiddles {
  DevBuf<cpx<double>> d[4];  // N0, N0/2, N1, N2
  DevBuf<cpx<float>> f[4];
};

void make_twiddles(Twiddles& tw, const int N[3], hipStream_t s) {
  const int len[4] = {N[0], N[0] / 2, N[1], N[2]};
  for (int i = 0; i < 4; ++i) {
    auto wd = twiddles<double>(len[i]);
    auto wf = twiddles<float>(len[i]);
    tw.d[i] = DevBuf<cpx<double>>(len[i]);
    tw.f[i] = DevBuf<cpx<float>>(len[i]);
    HIPCHK(hipMemcpyAsync(tw.d[i].p, wd.data(), wd.size() * sizeof(cpx<double>), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(tw.f[i].p, wf.data(), wf.size() * sizeof(cpx<float>), hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ DIIS solve
// Minimise |sum c_j r_j|^2 subject to sum c_j = 1 (Pulay).  G is the Gram
// matrix over the nh history entries (in order).  The bordered system
// [G/s -1; -1 0] [c; l] = [0; -1] (s the largest diagonal entry) is solved
// in FP64 by LU with partial pivoting and one step of iterative refinement
// against the unfactored matrix, on the stack: it runs for every solute of
// a batch between two residual passes, while the device waits.  Returns
// false if the system is singular or the result not finite.  mdiis_step
// solves the same system on the device with the same operations in the
// same order (sums of the back substitution from the last column down), and
// neither contracts products into fused multiply-adds, so both give the
// same coefficients to the bit.
bool diis_coefficients(const std::vector<double>& G, int nh, std::vector<double>& c) {
#pragma clang fp contract(off)
  constexpr int NM = MAX_DIIS + 1;
  const int n = nh + 1;
  if (nh < 1 || nh > MAX_DIIS) return false;
  double scale = 0.0;
  for (int i = 0; i < nh; ++i) scale = std::max(scale, std::fabs(G[i * nh + i]));
  if (!(scale > 0.0) || !std::isfinite(scale)) return false;
  double A0[NM][NM], A[NM][NM], b0[NM], x[NM], r[NM];
  int perm[NM];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      A0[i][j] = (i < nh && j < nh) ? G[i * nh + j] / scale : (i == nh && j == nh ? 0.0 : -1.0);
  for (int i = 0; i < n; ++i) {
    b0[i] = i == nh ? -1.0 : 0.0;
    perm[i] = i;
    for (int j = 0; j < n; ++j) A[i][j] = A0[i][j];
  }
  for (int col = 0; col < n; ++col) {  // PA = LU in place, L unit lower
    int piv = col;
    for (int rr = col + 1; rr < n; ++rr)
      if (std::fabs(A[rr][col]) > std::fabs(A[piv][col])) piv = rr;
    if (!(std::fabs(A[piv][col]) > 1e-300)) return false;
    if (piv != col) {
      for (int k = 0; k < n; ++k) std::swap(A[col][k], A[piv][k]);
      std::swap(perm[col], perm[piv]);
    }
    const double inv = 1.0 / A[col][col];
    for (int rr = col + 1; rr < n; ++rr) {
      const double f = A[rr][col] * inv;
      A[rr][col] = f;
      for (int k = col + 1; k < n; ++k) A[rr][k] -= f * A[col][k];
    }
  }
  auto lu_solve = [&](const double* rhs, double* out) {
    double y[NM];
    for (int i = 0; i < n; ++i) {
      double t = rhs[perm[i]];
      for (int k = 0; k < i; ++k) t -= A[i][k] * y[k];
      y[i] = t;
    }
    for (int i = n - 1; i >= 0; --i) {
      double t = y[i];
      for (int k = n - 1; k > i; --k) t -= A[i][k] * out[k];
      out[i] = t / A[i][i];
    }
  };
  lu_solve(b0, x);
  for (int i = 0; i < n; ++i) {  // one refinement step: x += A^-1 (b - A x)
    double t = b0[i];
    for (int k = 0; k < n; ++k) t -= A0[i][k] * x[k];
    r[i] = t;
  }
  double dx[NM];
  lu_solve(r, dx);
  c.assign(nh, 0.0);
  for (int i = 0; i < nh; ++i) {
    c[i] = x[i] + dx[i];
    if (!std::isfinite(c[i])) return false;
  }
  return true;
}

// --------------------------------------------------------------- solver
struct SoluteState {
  // device memory
  DevBuf<double> U;        // D slots of u_j = t_j + lam r_j, sized for FP64
  DevBuf<double> R;        // D slots of r_j
  DevBuf<double> t64;      // FP64 iterate
  DevBuf<float> t32;       // FP32 iterate (mixed only)
  DevBuf<double> bv64;     // beta v, FP64
  DevBuf<float> bv32;      // and FP32 (mixed only)
  DevBuf<cpx<double>> W;   // spectral work (FP32 phase uses the front half)
  DevBuf<cpx<double>> fs64;
  DevBuf<cpx<float>> fs32;  // y-major with the Nyquist column apart (fs_index)
  DevBuf<double> part;     // residual-workgroup partials, NPART each
  DevBuf<double> sites;    // n * 6
  DevBuf<double> Dt, gmu;  // adjoint: dc/dt and dmu/dt at the converged t
  DevBuf<double> r64;      // mixed: F(t_b) at a rebase, FP64
  DevBuf<float> fb32;      // mixed: F(t_b) for the FP32 correction passes
  DevBuf<float> t32b;      // mixed: second correction buffer
  float* dcur = nullptr;   // mixed: the correction delta of the current base
  float* dshift = nullptr; // mixed: the last delta folded into the base, by
                           // which entries of the previous base are moved
  float* ucur = nullptr;   // mixed, difference-form history: the newest u
  std::vector<float> rsc;  // mixed, FP8 residual history: the scale of each slot
  int base = 0;            // mixed: rebases so far in this FP32 phase
  std::vector<int> sbase;  // mixed: base of each history slot
  double tol_lin = 0.0;    // adjoint: absolute tolerance (relative to |g|)
  double tol_mixed = 0.0;  // mixed: FP64-checked max|F| that ends the FP32 phase
  int n_sites = 0;
  // host MDIIS state
  std::vector<int> order;     // history slots, oldest first
  std::vector<double> coef;   // t = sum coef_j u_{order_j} for the next step
  std::vector<int> dotted;    // slots the pending residual is dotted with
  int slot_new = 0;           // slot receiving the pending r and u
  std::vector<double> gram;   // D x D Gram matrix by slot
  int active = 1;
  int status = RISM3D_NONCONVERGED;
  int iters = 0, iters32 = 0;
  int phase32 = 0;            // 1 while the FP32 kernels run
  int reached_mixed = 0;      // FP32 phase ended at mixed_tol or its floor
  double best32 = INFINITY;   // FP32: best max|F| since the last rebase
  int stall32 = 0;
  double rb_rinf = INFINITY;  // FP32: max|F(t_b)| at the last rebase (FP64)
  int ready_at = -1;          // FP32: pass of the batch when ready for a rebase (-1: not)
  int parked = 0;             // FP32: at mixed_tol / 2, waiting for a rebase without passes
  int polish = 0;             // FP64 steps taken after the FP32 phase
  double resid = INFINITY, best = INFINITY;
  int since_best = 0;
};

struct Problem {
  GridArgs g;
  int N[3];
  double L[3], h[3];
  int D;          // DIIS depth (1 = Picard)
  int mixed;
  int r16 = env_int("RISM3D_R_BF16", 1);  // FP32 phase: residual history in BF16
  int D32 = env_int("RISM3D_DEPTH32", 0);   // FP32 phase: MDIIS depth (0: D)
  double lam32 = std::getenv("RISM3D_STEP32") ? std::atof(std::getenv("RISM3D_STEP32")) : 0.0;
  // FP32 phase (with r16): the u history in difference form, the newest u
  // in FP32 and the differences of consecutive entries in BF16
  int u16 = env_int("RISM3D_U_DIFF", 1);
  // FP32 phase (with r16): the residual history in FP8 e4m3, each entry
  // scaled by 240 / (max|r| of the previous pass) into the e4m3 range
  int r8 = env_int("RISM3D_R_FP8", 1);
  int B;
  int64_t mP;
  hipStream_t s;
  Twiddles tw;
  DevBuf<double> chi_tab, tau_tab;      // 1D tables [pairs][nrad], [m][nrad]
  DevBuf<double> chi3, tau3;            // on the 3D k grid, FP64
  DevBuf<float> chi3f, tau3f;           // and FP32
  DevBuf<SoluteDesc> d_sd;
  DevBuf<SitesDesc> d_sites;
  DevBuf<double> d_out;                 // B * NPART
  DevBuf<void*> d_cdst, d_csrc;         // precision hand-over copies
  DevBuf<int> d_cmask;
  int nparts = 0;
  std::vector<SoluteState> st;
  PinnedRing up;    // descriptor uploads
  PinnedRing down;  // norms and Gram rows
  // The residual chain (five kernels + the norm reduction) as HIP graphs,
  // one per working precision and history width, captured on first use and
  // replayed every iteration (the descriptors they read are uploaded
  // before each replay).  Re-captured if the grid arguments change.
  // RISM3D_GRAPHS=0 launches the kernels one by one instead.
  struct Graphs {
    // [FP64 / FP32 / FP32 with the device MDIIS step][history width 8 / 10 / 16]
    hipGraph_t graph[3][3] = {};
    hipGraphExec_t exec[3][3] = {};
    GridArgs key[3][3];
    ~Graphs() {
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          if (exec[a][b]) (void)hipGraphExecDestroy(exec[a][b]);
          if (graph[a][b]) (void)hipGraphDestroy(graph[a][b]);
        }
    }
  } graphs;
  int use_graphs = env_int("RISM3D_GRAPHS", 1);
  // RISM3D_DEVICE_MDIIS=1: the FP32 phase's MDIIS step on the device
  // (mdiis_step, with the difference-form u and FP8 r histories): its own
  // descriptors, which the step rewrites, the solutes' history state and
  // the pinned status array the host polls after each pass.  Off by
  // default: the same step on the host measured about 1% faster, on the
  // 64-solute batch and on the natural grids alike.
  int dev_mdiis = env_int("RISM3D_DEVICE_MDIIS", 0);
  // FP32 phase: a solute whose residual fell 1e3 below its base's waits for
  // a rebase, iterating on, until another solute needs one or for at most
  // this many passes, so that rebases run together (0: each at once)
  int rebase_wait = env_int("RISM3D_REBASE_WAIT", 8);
  DevBuf<SoluteDesc> d_sd32;
  DevBuf<MdiisDev> d_mdiis;
  struct PinnedStatus {
    MdiisStatus* p = nullptr;
    ~PinnedStatus() {
      if (p) (void)hipHostFree(p);
    }
  } h_status;
  // launch configuration of the residual chain per working precision
  // ([0] FP64, [1] FP32): defaults from the environment, or the autotuner's
  // choice (tune)
  struct Knobs {
    int planes;  // FP32 chain on xy planes (use_planes)
    int zorder;  // z-pass block order (zstar_reg)
    int zwide;   // z-pass tile of 32 kx columns over 512 threads (FP32)
    int nyq;     // Nyquist column in a launch of its own (FP32)
  } knob[2] = {{0, env_int("RISM3D_ZSTAR_ORDER", 3), 0, 0},
               {env_int("RISM3D_PLANES", 1), env_int("RISM3D_ZSTAR_ORDER", 3),
                env_int("RISM3D_ZSTAR_WIDE", 0), env_int("RISM3D_NYQ", 1)}};
  int noalloc = 0;  // buffers from the fixed scratch: no pinned staging, no graphs
};

// Descriptors to the device through pinned staging (stream-ordered).
void upload_sd(Problem& pb, const std::vector<SoluteDesc>& sd, SoluteDesc* dst) {
  const size_t bytes = sd.size() * sizeof(SoluteDesc);
  if (pb.noalloc) {  // from pageable memory, complete before sd can change
    HIPCHK(hipMemcpyAsync(dst, sd.data(), bytes, hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
    return;
  }
  int k = 0;
  void* h = pb.up.slot(bytes, &k);
  std::memcpy(h, sd.data(), bytes);
  HIPCHK(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, pb.s));
  pb.up.mark(k, pb.s);
}
// pb.d_out to out[], synchronising the stream.
void download_out(Problem& pb, std::vector<double>& out) {
  const size_t bytes = out.size() * sizeof(double);
  int k = 0;
  void* h = pb.down.slot(bytes, &k);
  HIPCHK(hipMemcpyAsync(h, pb.d_out.p, bytes, hipMemcpyDeviceToHost, pb.s));
  pb.down.mark(k, pb.s);
  HIPCHK(hipStreamSynchronize(pb.s));
  std::memcpy(out.data(), h, bytes);
}

// Per-kernel HIP-event timing of the residual chain (rism3d_profile_*).
// Kernel ids: 0 closure_r2c_x, 1 fft_y (forward), 2 zstar, 3 fft_y
// (inverse), 4 c2r_x_resid, 5 closure_r2c_xy, 6 c2r_xy_resid; slots
// 0..PROF_KERNELS-1 count the FP64 launches, the next PROF_KERNELS the
// FP32 ones.  Bytes are the algorithmic (compulsory) HBM traffic of each
// launch: every array the kernel must read or write, once.
constexpr int PROF_KERNELS = 8;
struct Profile {
  bool on = false;
  hipEvent_t ev[5][2] = {};
  double ms[2 * PROF_KERNELS] = {};
  double bytes[2 * PROF_KERNELS] = {};
  long launches[2 * PROF_KERNELS] = {};
};
Profile g_prof;

int resid_grid_x(const GridArgs& g) {
  const int rows = rows_for(g.N0 / 2);
  return (g.N1 * g.N2 + rows - 1) / rows;
}

// History slots the closure and residual kernels must unroll for this batch.
int hist_of(const std::vector<SoluteDesc>& sd) {
  int h = 0;
  for (const auto& d : sd)
    if (d.active) h = std::max(h, std::max(d.ncomb, d.ndot));
  return h;
}

// One residual pass over the batch.  step 0: the host's descriptors sd.
// step 1 (FP32): the descriptors already in pb.d_sd32, with hist2 history
// slots at most; the device MDIIS step (mdiis_step) follows the pass and
// writes the next pass's descriptors there.
void evaluate(Problem& pb, bool fp32, std::vector<SoluteDesc>& sd, int step = 0, int hist2 = 0) {
  GridArgs g = pb.g;
  const Problem::Knobs& kn = pb.knob[fp32 ? 1 : 0];
  g.zorder = std::min(3, std::max(0, kn.zorder));
  g.zwide = kn.zwide;
  g.nyq = kn.nyq;
  SoluteDesc* dsd = step ? pb.d_sd32.p : pb.d_sd.p;
  Profile& pf = g_prof;
  std::vector<SoluteDesc> sdp;  // profiling, step 2: the descriptors on the device
  if (!step) {
    upload_sd(pb, sd, dsd);
  } else if (pf.on) {
    sdp.resize(sd.size());
    HIPCHK(hipMemcpyAsync(sdp.data(), dsd, sdp.size() * sizeof(SoluteDesc), hipMemcpyDeviceToHost,
                          pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
  }
  const std::vector<SoluteDesc>& sdb = sdp.empty() ? sd : sdp;
  const dim3 gx(resid_grid_x(g), g.m, pb.B);
  auto mark = [&](int k, int which) {
    if (pf.on) HIPCHK(hipEventRecord(pf.ev[k][which], pb.s));
  };
  const int hist = step ? hist2 : hist_of(sd);
  const bool planes = use_planes(g, fp32, kn.planes);
  if (planes) {
    // y-major spectra, the Nyquist column in the z pass's launch of its
    // own (32-column tiles only where they end at it)
    g.ylay = 1;
    g.nyq = 1;
    if (g.zwide && (g.NX - 1) % 32 != 0) g.zwide = 0;
  }
  const int nparts = planes ? g.m * g.N2 : pb.nparts;  // partial rows per solute
  auto chain = [&]() {
    if (planes) {
      using T = float;
      mark(0, 0); launch_planes<T>(false, pb.B, pb.s, dsd, g, pb.tw.f[1].p, pb.tw.f[0].p, pb.tw.f[2].p, hist); mark(0, 1);
      mark(2, 0); launch_zstar<T>(g.N2, g.NX, g.N1, pb.B, pb.s, dsd, g, pb.tw.f[3].p, pb.chi3f.p, pb.tau3f.p); mark(2, 1);
      mark(4, 0); launch_planes<T>(true, pb.B, pb.s, dsd, g, pb.tw.f[1].p, pb.tw.f[0].p, pb.tw.f[2].p, hist); mark(4, 1);
    } else if (fp32) {
      using T = float;
      constexpr int LN = y_lines<T>();
      const dim3 gy((g.NX + LN - 1) / LN, g.m * g.N2, pb.B);
      mark(0, 0); launch_closure_r2c<T>(g.N0, gx, pb.s, dsd, g, pb.tw.f[1].p, pb.tw.f[0].p, hist); mark(0, 1);
      mark(1, 0); launch_fft_y<T, false>(g.N1, gy, pb.s, dsd, g, pb.tw.f[2].p); mark(1, 1);
      mark(2, 0); launch_zstar<T>(g.N2, g.NX, g.N1, pb.B, pb.s, dsd, g, pb.tw.f[3].p, pb.chi3f.p, pb.tau3f.p); mark(2, 1);
      mark(3, 0); launch_fft_y<T, true>(g.N1, gy, pb.s, dsd, g, pb.tw.f[2].p); mark(3, 1);
      mark(4, 0); launch_c2r_resid<T>(g.N0, gx, pb.s, dsd, g, pb.tw.f[1].p, pb.tw.f[0].p, hist); mark(4, 1);
    } else {
      using T = double;
      constexpr int LN = y_lines<T>();
      const dim3 gy((g.NX + LN - 1) / LN, g.m * g.N2, pb.B);
      mark(0, 0); launch_closure_r2c<T>(g.N0, gx, pb.s, dsd, g, pb.tw.d[1].p, pb.tw.d[0].p, hist); mark(0, 1);
      mark(1, 0); launch_fft_y<T, false>(g.N1, gy, pb.s, dsd, g, pb.tw.d[2].p); mark(1, 1);
      mark(2, 0); launch_zstar<T>(g.N2, g.NX, g.N1, pb.B, pb.s, dsd, g, pb.tw.d[3].p, pb.chi3.p, pb.tau3.p); mark(2, 1);
      mark(3, 0); launch_fft_y<T, true>(g.N1, gy, pb.s, dsd, g, pb.tw.d[2].p); mark(3, 1);
      mark(4, 0); launch_c2r_resid<T>(g.N0, gx, pb.s, dsd, g, pb.tw.d[1].p, pb.tw.d[0].p, hist); mark(4, 1);
    }
    hipLaunchKernelGGL(reduce_resid, dim3(pb.B), dim3(NT), 0, pb.s, dsd, nparts, pb.d_out.p);
    if (step)
      hipLaunchKernelGGL(mdiis_step, dim3(pb.B), dim3(64), 0, pb.s, pb.d_out.p, pb.d_mdiis.p, dsd,
                         pb.h_status.p);
  };
  if (pb.use_graphs && !pf.on) {
    auto& G = pb.graphs;
    const int a = step ? 2 : (fp32 ? 1 : 0), b = hist <= 8 ? 0 : (hist <= 10 ? 1 : 2);
    if (!G.exec[a][b] || std::memcmp(&G.key[a][b], &g, sizeof g) != 0) {
      if (G.exec[a][b]) HIPCHK(hipGraphExecDestroy(G.exec[a][b]));
      if (G.graph[a][b]) HIPCHK(hipGraphDestroy(G.graph[a][b]));
      G.exec[a][b] = nullptr;
      G.graph[a][b] = nullptr;
      HIPCHK(hipStreamBeginCapture(pb.s, hipStreamCaptureModeThreadLocal));
      chain();
      HIPCHK(hipStreamEndCapture(pb.s, &G.graph[a][b]));
      HIPCHK(hipGraphInstantiate(&G.exec[a][b], G.graph[a][b], nullptr, nullptr, 0));
      std::memcpy(&G.key[a][b], &g, sizeof g);
    }
    HIPCHK(hipGraphLaunch(G.exec[a][b], pb.s));
  } else {
    chain();
  }
  if (pf.on) {
    const double sc = fp32 ? 4.0 : 8.0;  // working-type scalar
    const double Kv = (double)g.NX * g.N1 * g.N2;  // spectral points without padding
    const double mP = (double)g.m * g.P, mK = (double)g.m * Kv;
    double b0 = 0, b4 = 0;
    int na = 0;
    for (auto& d : sdb) {
      if (!d.active) continue;
      ++na;
      // u_j in (+ the shift of the previous base's entries), t out (or t
      // in), beta v in, C out
      const double uin = d.u16 ? mP * sc + (d.ncomb - 1) * mP * 2.0 : d.ncomb * mP * sc;
      b0 += uin + (d.shift_w != 0.0 ? mP * sc : 0) + mP * sc + mP * sc + mK * 2 * sc;
      // W in, t in, r and u out, r_j in (+ F(t_b) in, defect correction);
      // r and r_j in BF16 with r16
      const double sr = d.r8 ? 1.0 : (d.r16 ? 2.0 : sc);
      b4 += mK * 2 * sc + mP * sc + mP * sr + mP * sc + d.ndot * mP * sr +
            (d.mode == 5 ? mP * sc : 0) + (d.u16 ? mP * sc + mP * 2.0 : 0);
    }
    const double tables = (double)(g.m * (g.m + 1) / 2 + g.m) * Kv * sc;
    const double per[5] = {b0, na * 2 * mK * 2 * sc,
                           na * (2 * mK * 2 * sc + Kv * 2 * sc) + tables,
                           na * 2 * mK * 2 * sc, b4};
    HIPCHK(hipEventSynchronize(pf.ev[4][1]));
    const int slot = fp32 ? PROF_KERNELS : 0;
    for (int k = 0; k < 5; ++k) {
      if (planes && (k == 1 || k == 3)) continue;  // the y passes are inside the plane kernels
      const int id = planes ? (k == 0 ? 5 : k == 4 ? 6 : k) : k;
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, pf.ev[k][0], pf.ev[k][1]));
      pf.ms[slot + id] += ms;
      pf.bytes[slot + id] += per[k];
      pf.launches[slot + id] += 1;
    }
  }
}

// ------------------------------------------------------------ autotuner
// Candidate launch configurations of the residual chain, timed on the
// batch's own buffers: five evaluations of F at t = 0 for every solute
// after one untimed, launched one by one, HIP events on the solver's
// stream.  The first candidate is the built-in configuration; another
// replaces it only when at least 1% faster (closer choices are within the
// trial noise).  The choice is kept in pb.knob for this working precision.
// Choices are cached per process by grid, sites, batch size and
// precision, loaded from and appended to RISM3D_TUNE_FILE if it is set.
struct TuneTable {
  std::mutex mu;
  std::map<std::string, std::pair<Problem::Knobs, double>> best;
  bool loaded = false;
};
TuneTable& tune_table() {
  static TuneTable t;
  return t;
}
std::string tune_key(const GridArgs& g, int B, bool fp32) {
  char buf[96];
  std::snprintf(buf, sizeof buf, "%dx%dx%d,m%d,B%d,%s", g.N0, g.N1, g.N2, g.m, B,
                fp32 ? "fp32" : "fp64");
  return buf;
}
void tune(Problem& pb, bool fp32, int verbose) {
  const GridArgs& g0 = pb.g;
  const std::string key = tune_key(g0, pb.B, fp32);
  TuneTable& T = tune_table();
  const char* file = std::getenv("RISM3D_TUNE_FILE");
  {
    std::lock_guard<std::mutex> lk(T.mu);
    if (!T.loaded && file) {  // lines: key planes zorder zwide nyq ms
      if (FILE* f = std::fopen(file, "r")) {
        char k[96];
        Problem::Knobs c;
        double ms;
        while (std::fscanf(f, "%95s %d %d %d %d %lf", k, &c.planes, &c.zorder, &c.zwide, &c.nyq,
                           &ms) == 6)
          T.best[k] = {c, ms};
        std::fclose(f);
      }
    }
    T.loaded = true;
    auto it = T.best.find(key);
    if (it != T.best.end()) {
      pb.knob[fp32 ? 1 : 0] = it->second.first;
      return;
    }
  }
  std::vector<Problem::Knobs> cand{pb.knob[fp32 ? 1 : 0]};  // the built-in one first
  const bool can_planes = fp32 && use_planes(g0, true, 1);
  const bool can_wide = fp32 && g0.N2 <= 128 && use_zreg(g0);
  for (int planes : {1, 0})
    for (int zorder : {3, 2, 1})
      for (int zwide : {0, 1})
        for (int nyq : {1, 0}) {
          if ((planes && !can_planes) || (zwide && !can_wide) || (!fp32 && (nyq || zwide)) ||
              (planes && !nyq))  // the plane chain always splits the Nyquist column off
            continue;
          const Problem::Knobs& d0 = cand[0];
          if (planes == d0.planes && zorder == d0.zorder && zwide == d0.zwide && nyq == d0.nyq)
            continue;
          cand.push_back({planes, zorder, zwide, nyq});
        }
  // F(t) at t = 0 for every solute, into history slot 0
  std::vector<SoluteDesc> sd(pb.B);
  for (int b = 0; b < pb.B; ++b) {
    auto& s = pb.st[b];
    SoluteDesc& d = sd[b];
    std::memset(&d, 0, sizeof d);
    d.active = 1;
    if (fp32) HIPCHK(hipMemsetAsync(s.t32.p, 0, pb.mP * 4, pb.s));
    d.t = fp32 ? (void*)s.t32.p : (void*)s.t64.p;
    d.r = s.R.p;
    d.unew = s.U.p;
    d.lam = 0.5;
    d.bv = fp32 ? (const void*)s.bv32.p : (const void*)s.bv64.p;
    d.W = s.W.p;
    d.Fs = fp32 ? (const void*)s.fs32.p : (const void*)s.fs64.p;
    d.part = s.part.p;
  }
  struct Pause {  // eager launches, and no kernel timing of the trials
    Problem& pb;
    int graphs;
    bool prof;
    explicit Pause(Problem& p) : pb(p), graphs(p.use_graphs), prof(g_prof.on) {
      pb.use_graphs = 0;
      g_prof.on = false;
    }
    ~Pause() {
      pb.use_graphs = graphs;
      g_prof.on = prof;
    }
  } pause(pb);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  Problem::Knobs best = pb.knob[fp32 ? 1 : 0];
  double best_ms = INFINITY;
  for (const auto& c : cand) {
    pb.knob[fp32 ? 1 : 0] = c;
    evaluate(pb, fp32, sd);  // untimed: first launch of these kernels
    HIPCHK(hipEventRecord(e0, pb.s));
    for (int r = 0; r < 5; ++r) evaluate(pb, fp32, sd);
    HIPCHK(hipEventRecord(e1, pb.s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    ms /= 5;
    if (verbose)
      std::fprintf(stderr, "rism3d: tune %s planes %d zorder %d zwide %d nyq %d: %.3f ms\n",
                   key.c_str(), c.planes, c.zorder, c.zwide, c.nyq, ms);
    if (ms < (best_ms == INFINITY ? INFINITY : 0.99 * best_ms)) { best_ms = ms; best = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  pb.knob[fp32 ? 1 : 0] = best;
  std::lock_guard<std::mutex> lk(T.mu);
  T.best[key] = {best, best_ms};
  if (file)
    if (FILE* f = std::fopen(file, "a")) {
      std::fprintf(f, "%s %d %d %d %d %.4f\n", key.c_str(), best.planes, best.zorder, best.zwide,
                   best.nyq, best_ms);
      std::fclose(f);
    }
}

struct Config {
  int solver, precision, polish, depth, verbose, max_iter;
  double lam, tol, mixed_tol;
  int mode = 0;          // 0: F(t) = 0; 1: adjoint lambda = D K lambda + g
  double rel_tol = 0.0;  // mode 1: stop at max|r| <= rel_tol max|g|
};

// Copy the iterate between its FP32 and FP64 buffers for the solutes whose
// mask entry is set.
void convert_iterates(Problem& pb, const std::vector<int>& mask, bool to64) {
  const int B = pb.B;
  std::vector<void*> dst(B), src(B);
  for (int b = 0; b < B; ++b) {
    dst[b] = to64 ? (void*)pb.st[b].t64.p : (void*)pb.st[b].t32.p;
    src[b] = to64 ? (void*)pb.st[b].t32.p : (void*)pb.st[b].t64.p;
  }
  HIPCHK(hipMemcpyAsync(pb.d_cdst.p, dst.data(), B * sizeof(void*), hipMemcpyHostToDevice, pb.s));
  HIPCHK(hipMemcpyAsync(pb.d_csrc.p, src.data(), B * sizeof(void*), hipMemcpyHostToDevice, pb.s));
  HIPCHK(hipMemcpyAsync(pb.d_cmask.p, mask.data(), B * sizeof(int), hipMemcpyHostToDevice, pb.s));
  const int ch = (int)std::min<int64_t>(2048, (pb.mP + NT - 1) / NT);
  hipLaunchKernelGGL(convert_batch, dim3(ch, B), dim3(NT), 0, pb.s, pb.d_cdst.p,
                     (const void* const*)pb.d_csrc.p, to64 ? 1 : 0, pb.mP, pb.d_cmask.p);
  HIPCHK(hipStreamSynchronize(pb.s));
}

// Defect correction (mixed precision).  The FP32 passes iterate on a
// correction delta around an FP64 base t_b: F(t_b + delta) = F(t_b) +
// K (c(t_b + delta) - c(t_b)) - delta, with F(t_b) from one FP64 residual
// evaluation and the closure difference formed from x_b = -beta v + t_b in
// a form whose FP32 rounding is relative to |delta c| (closure_dc).  The
// FP32 error then scales with the size of the correction, not with |t|
// and |c| (hundreds near charged sites, where a plain FP32 iteration stalls
// at max|F| ~ 1e-4).  A rebase folds delta into t_b (FP64), evaluates
// F(t_b) in FP64 -- the residual the stopping test uses -- and restarts
// the FP32 MDIIS history with r_0 = F(t_b).  add = 0 on entry (delta = 0).
// Returns, per solute of the mask, the FP64 max|F(t_b)| and ||F(t_b)||^2.
void rebase(Problem& pb, const Config& cf, const std::vector<int>& mask, bool add,
            std::vector<double>& rinf, std::vector<double>& r2) {
  const int B = pb.B;
  const int64_t mP = pb.mP;
  const int ch = (int)std::min<int64_t>(2048, (mP + NT - 1) / NT);
  std::vector<SoluteDesc> sd(B);
  for (int b = 0; b < B; ++b) {
    SoluteDesc& d = sd[b];
    std::memset(&d, 0, sizeof d);
    if (!mask[b]) continue;
    auto& s = pb.st[b];
    if (add) hipLaunchKernelGGL(add_delta, dim3(ch), dim3(NT), 0, pb.s, s.t64.p, s.dcur, mP);
    d.active = 1;
    d.t = s.t64.p;
    d.r = s.r64.p;
    d.unew = nullptr;  // the FP32 history lives in U: no u written
    d.lam = cf.lam;
    d.bv = s.bv64.p;
    d.W = s.W.p;
    d.Fs = s.fs64.p;
    d.part = s.part.p;
    if (cf.mode == 1) {  // the adjoint's residual D K l + g - l
      d.mode = 1;
      d.dt = s.Dt.p;
      d.gv = s.gmu.p;
    }
  }
  evaluate(pb, false, sd);
  std::vector<double> out((size_t)B * NPART);
  download_out(pb, out);
  HIPCHK(hipGetLastError());
  rinf.assign(B, 0.0);
  r2.assign(B, 0.0);
  for (int b = 0; b < B; ++b) {
    if (!mask[b]) continue;
    rinf[b] = out[(size_t)b * NPART];
    r2[b] = out[(size_t)b * NPART + 1];
  }
}

// New base after rebase(): x_b and F(t_b) for one solute.  seed (the first
// base of the FP32 phase): delta = 0 and the first MDIIS entry, r_0 =
// F(t_b), u_0 = lam r_0 in slot 0.  Later bases keep the MDIIS history: the
// delta just folded into t_b becomes the shift of the entries of the old
// base (t = sum_j c_j u_j - (sum_{old j} c_j) shift), entries of the base
// before are dropped, and the coefficients are refitted if any was.

// The FP32 phase's MDIIS step (RISM3D_STEP32, default the solve's step).
double lam32(const Problem& pb, const Config& cf) { return pb.lam32 > 0 ? pb.lam32 : cf.lam; }

void rebase_start(Problem& pb, const Config& cf, int b, double r2, bool seed) {
  auto& s = pb.st[b];
  s.ready_at = -1;
  s.parked = 0;
  const int64_t mP = pb.mP;
  const int D = pb.D;
  const int ch = (int)std::min<int64_t>(2048, (mP + NT - 1) / NT);
  float* U0 = reinterpret_cast<float*>(s.U.p);
  void* R0 = s.R.p;  // slot 0, BF16 or FP32
  if (seed) {
    s.dcur = s.t32.p;
    s.dshift = s.t32b.p;
    s.base = 0;
    s.sbase.assign(D, 0);
    // difference form: D BF16 slots at the front of U, the newest u behind
    s.ucur = pb.u16 ? reinterpret_cast<float*>(reinterpret_cast<char*>(s.U.p) +
                                               (((size_t)D * mP * 2 + 255) & ~(size_t)255))
                    : nullptr;
    if (pb.u16) U0 = s.ucur;
  }
  float r8s = 0.f;  // FP8 history: r_0 = F(t_b) scaled by 240 / max|F(t_b)|
  if (seed && pb.r8) {
    s.rsc.assign(D, 1.f);
    r8s = fp8_scale(s.resid);
    s.rsc[0] = r8s;
  }
  hipLaunchKernelGGL(rebase_prep, dim3(ch), dim3(NT), 0, pb.s, s.t64.p, s.bv64.p, s.r64.p,
                     pb.g.closure == RISM3D_CLOSURE_PY ? 1 : 0, seed ? 1 : 0, (float)lam32(pb, cf),
                     s.bv32.p, s.fb32.p, s.dcur, U0, R0, pb.r16, r8s, mP);
  s.best32 = INFINITY;
  s.stall32 = 0;
  if (seed) {
    s.order.assign(1, 0);
    s.coef.assign(1, 1.0);
    s.gram.assign((size_t)D * D, 0.0);
    s.gram[0] = r2;
    s.best = r2;
    s.since_best = 0;
    return;
  }
  std::vector<int> keep;
  for (int q : s.order)
    if (s.sbase[q] == s.base) keep.push_back(q);
  if (keep.size() != s.order.size()) {
    s.order = keep;
    const int nh = (int)s.order.size();
    std::vector<double> G(nh * nh);
    for (int i = 0; i < nh; ++i)
      for (int j = 0; j < nh; ++j) G[i * nh + j] = s.gram[s.order[i] * D + s.order[j]];
    if (nh == 1 || !diis_coefficients(G, nh, s.coef)) {
      s.order.assign(1, s.order.back());
      s.coef.assign(1, 1.0);
    }
  }
  std::swap(s.dcur, s.dshift);
  s.base++;
  // difference form: the newest u moves to the new base (differences of
  // entries do not change), so no shift enters the later steps
  if (pb.u16) hipLaunchKernelGGL(shift_u, dim3(ch), dim3(NT), 0, pb.s, s.ucur, s.dshift, mP);
}

// Runs the fixed-point iteration for every solute of the batch.  MDIIS
// (Pulay) over the history u_j = t_j + lam r_j: t = sum c_j u_j with c
// minimising |sum c_j r_j|, sum c_j = 1.  Depth 1 is Picard, t += lam r
// (bgy3d_snes_picard, bgy3d-snes.c:348-384).  The host reads the norms and
// the Gram row back once per iteration and takes the step itself; with
// RISM3D_DEVICE_MDIIS=1 the FP32 phase of a mixed solve takes it on the
// device (mdiis_step) and the host polls a pinned status per pass.
void iterate_batch(Problem& pb, const Config& cf) {
  const int B = pb.B;
  const int64_t mP = pb.mP;
  const int D = pb.D;
  std::vector<SoluteDesc> sd(B);
  std::vector<double> out((size_t)B * NPART);
  std::vector<int> act(B, 0);
  for (int b = 0; b < B; ++b) {
    auto& s = pb.st[b];
    act[b] = s.active != 0;
    s.phase32 = s.active && cf.precision != RISM3D_FP64;
    s.gram.assign(D * D, 0.0);
    s.order.clear();
    s.coef.clear();
  }
  // mixed: FP64 residual at the starting point, the first base
  auto end_fp32 = [&](int b, double rinf, double r2) {
    // the base passed the FP64 test at mixed_tol: polish or stop
    auto& s = pb.st[b];
    s.phase32 = 0;
    s.reached_mixed = 1;
    s.resid = rinf;
    s.polish = 0;
    if (rinf <= cf.tol || cf.polish == 0) {
      s.status = RISM3D_OK;
      s.active = 0;
      return;
    }
    // FP64 history entry 0 from the rebase: r_0 = F(t_b), u_0 = t_b + lam r_0
    HIPCHK(hipMemcpyAsync(s.R.p, s.r64.p, mP * 8, hipMemcpyDeviceToDevice, pb.s));
    const int ch = (int)std::min<int64_t>(2048, (mP + NT - 1) / NT);
    hipLaunchKernelGGL(seed_u, dim3(ch), dim3(NT), 0, pb.s, s.U.p, s.t64.p, s.r64.p, cf.lam, mP);
    s.order.assign(1, 0);
    s.coef.assign(1, 1.0);
    s.gram.assign((size_t)D * D, 0.0);
    s.gram[0] = r2;
    s.best = r2;
    s.since_best = 0;
  };
  auto after_rebase = [&](const std::vector<int>& mask, const std::vector<double>& rinf,
                          const std::vector<double>& r2, bool seed) {
    for (int b = 0; b < B; ++b) {
      if (!mask[b]) continue;
      auto& s = pb.st[b];
      s.parked = 0;
      s.ready_at = -1;
      s.iters++;
      s.resid = rinf[b];
      if (cf.verbose)
        std::fprintf(stderr, "rism3d: solute %d iter %d fp64 rebase |F|inf = %.3e\n", b, s.iters,
                     rinf[b]);
      if (!std::isfinite(rinf[b]) || !std::isfinite(r2[b])) {
        s.status = RISM3D_NAN;
        s.active = 0;
        s.phase32 = 0;
        continue;
      }
      if (seed) {
        // the adjoint starts at l = 0, where the residual is g: its
        // tolerance is relative to max|g|
        s.tol_mixed = cf.mode == 1 ? cf.rel_tol * rinf[b] : cf.mixed_tol;
        if (cf.mode == 1) s.tol_lin = s.tol_mixed;
      }
      if (rinf[b] <= s.tol_mixed || rinf[b] <= cf.tol) { end_fp32(b, rinf[b], r2[b]); continue; }
      if (cf.precision == RISM3D_FP32 && !seed) {
        // FP32 only: no second base; the FP64 residual at the end is
        // reported, OK if it reached mixed_tol
        s.status = RISM3D_NONCONVERGED;
        s.active = 0;
        s.phase32 = 0;
        continue;
      }
      if (s.iters32 >= cf.max_iter) {
        s.status = RISM3D_NONCONVERGED;
        s.active = 0;
        s.phase32 = 0;
        continue;
      }
      s.rb_rinf = rinf[b];
      rebase_start(pb, cf, b, r2[b], seed);
    }
  };
  if (cf.precision != RISM3D_FP64) {
    std::vector<double> rinf, r2;
    rebase(pb, cf, act, false, rinf, r2);
    after_rebase(act, rinf, r2, true);
  }
  // the FP32 phase's MDIIS step on the device (mdiis_step): the history in
  // difference form with FP8 residuals, the library's own allocations
  const bool dev_ok = pb.dev_mdiis && pb.u16 && pb.r8 && pb.r16 && !pb.noalloc &&
                      cf.precision != RISM3D_FP64;
  bool dev_on = false;          // the FP32 solutes' state is on the device
  std::vector<int> dev_dirty;   // solutes whose host state goes up before the next pass
  std::vector<int> next_hist(B, 0);  // history slots of each solute's next pass
  std::vector<MdiisDev> hm(B);
  // host state of solute b (after make_desc) to its device form, and back
  auto pack = [&](int b, int Dp) {
    const auto& s = pb.st[b];
    MdiisDev& m = hm[b];
    std::memset(&m, 0, sizeof m);
    m.ucur = s.ucur;
    m.U = static_cast<uint16_t*>(static_cast<void*>(s.U.p));
    m.R = static_cast<uint8_t*>(static_cast<void*>(s.R.p));
    m.mP = mP;
    m.rb_rinf = s.rb_rinf;
    m.tol_half = 0.5 * s.tol_mixed;
    m.D = Dp;
    m.only32 = cf.precision == RISM3D_FP32;
    m.max_iter = cf.max_iter;
    m.base = s.base;
    m.n = (int)s.order.size();
    m.slot_new = s.slot_new;
    m.ndot = (int)s.dotted.size();
    m.since_best = s.since_best;
    m.stall32 = s.stall32;
    m.iters = s.iters;
    m.iters32 = s.iters32;
    m.hold = pb.rebase_wait > 0;
    for (int i = 0; i < m.n; ++i) {
      m.order[i] = s.order[i];
      m.coef[i] = s.coef[i];
    }
    for (int i = 0; i < m.ndot; ++i) m.dotted[i] = s.dotted[i];
    for (int q = 0; q < D; ++q) {
      m.sbase[q] = s.sbase[q];
      m.rsc[q] = s.rsc[q];
      for (int j = 0; j < D; ++j) m.gram[q * MAX_DIIS + j] = s.gram[q * D + j];
    }
    m.resid = s.resid;
    m.best = s.best;
    m.best32 = s.best32;
  };
  auto unpack = [&](int b, const MdiisDev& m) {
    auto& s = pb.st[b];
    s.order.assign(m.order, m.order + m.n);
    s.coef.assign(m.coef, m.coef + m.n);
    s.dotted.assign(m.dotted, m.dotted + m.ndot);
    s.slot_new = m.slot_new;
    for (int q = 0; q < D; ++q) {
      s.sbase[q] = m.sbase[q];
      s.rsc[q] = m.rsc[q];
      for (int j = 0; j < D; ++j) s.gram[q * D + j] = m.gram[q * MAX_DIIS + j];
    }
    s.since_best = m.since_best;
    s.stall32 = m.stall32;
    s.iters = m.iters;
    s.iters32 = m.iters32;
    s.resid = m.resid;
    s.best = m.best;
    s.best32 = m.best32;
  };
  // the host's descriptors and device state of the given solutes into
  // pb.d_sd32 and pb.d_mdiis: one launch of mdiis_load over pinned rows
  auto load_rows = [&](const std::vector<int>& rows) {
    constexpr size_t WS = sizeof(SoluteDesc) / 8, WM = sizeof(MdiisDev) / 8, W = 1 + WS + WM;
    int k = 0;
    uint64_t* h = static_cast<uint64_t*>(pb.up.slot(rows.size() * W * 8, &k));
    for (size_t i = 0; i < rows.size(); ++i) {
      const int b = rows[i];
      h[i * W] = (uint64_t)b;
      std::memcpy(h + i * W + 1, &sd[b], sizeof(SoluteDesc));
      std::memcpy(h + i * W + 1 + WS, &hm[b], sizeof(MdiisDev));
    }
    hipLaunchKernelGGL(mdiis_load, dim3((unsigned)rows.size()), dim3(64), 0, pb.s, h,
                       pb.d_sd32.p, pb.d_mdiis.p);
    pb.up.mark(k, pb.s);
  };
  // The end of a pass with the device step: every solute of the pass has
  // released its status with the next iteration count.  The host polls the
  // pinned statuses instead of waiting for the stream (which wakes it
  // later); a stream that ends or fails without them is an error.
  auto wait_status = [&]() {
    std::vector<int> expect(B, 0);
    for (int b = 0; b < B; ++b)
      if (pb.st[b].active && pb.st[b].phase32 && !pb.st[b].parked) expect[b] = pb.st[b].iters + 1;
    auto arrived = [&]() {
      for (int b = 0; b < B; ++b)
        if (expect[b] && __atomic_load_n(&pb.h_status.p[b].iters, __ATOMIC_ACQUIRE) != expect[b])
          return false;
      return true;
    };
    // the stream is asked only now and then (a failed pass ends the wait
    // within milliseconds): its lock is shared with the launches of other
    // threads, such as the size classes of rism3d_batch_solve_grids
    for (long spin = 1;; ++spin) {
      if (arrived()) break;
      __builtin_ia32_pause();
      if (spin % 16384 == 0) {
        const hipError_t e = hipStreamQuery(pb.s);
        if (e == hipSuccess) {
          if (arrived()) break;
          std::fprintf(stderr, "rism3d: mdiis_step: the pass ended without its statuses\n");
          throw HipFailure{hipErrorUnknown};
        }
        if (e != hipErrorNotReady) HIPCHK(e);
      }
    }
    HIPCHK(hipGetLastError());
  };
  // Rebases together: once a solute must rebase, or one ready has waited
  // rebase_wait passes, or every solute of the phase is ready, every solute
  // ready for one joins it.
  int npass32 = 0;  // FP32 passes of the batch so far
  auto group_rebases = [&](std::vector<int>& mask, bool& any_rebase, std::vector<int>* back) {
    if (pb.rebase_wait <= 0) return;
    bool go = any_rebase, all = true, some = false;
    for (int b = 0; b < B; ++b) {
      const auto& s = pb.st[b];
      if (!s.active || !s.phase32 || mask[b]) continue;
      if (s.ready_at < 0) { all = false; continue; }
      some = true;
      if (npass32 - s.ready_at >= pb.rebase_wait) go = true;
    }
    if (!go && !(all && some)) return;
    for (int b = 0; b < B; ++b) {
      const auto& s = pb.st[b];
      if (s.active && s.phase32 && !mask[b] && s.ready_at >= 0) {
        mask[b] = 1;
        any_rebase = true;
        if (back) back->push_back(b);
      }
    }
  };
  for (;;) {
    bool any = false, any32 = false;
    for (int b = 0; b < B; ++b) {
      auto& s = pb.st[b];
      if (s.active) { any = true; any32 |= (s.phase32 != 0); }
    }
    if (!any) break;
    // All solutes of a launch run one precision: FP32 while any still needs
    // FP32 (the others wait), then FP64.
    const bool fp32 = any32;
    const size_t sz = fp32 ? 4 : 8;
    const int Dp = fp32 && pb.D32 > 0 ? std::min(D, pb.D32) : D;  // this phase's depth
    // the descriptor of solute b's next pass from its host state
    auto make_desc = [&](int b) {
      auto& s = pb.st[b];
      SoluteDesc& d = sd[b];
      std::memset(&d, 0, sizeof d);
      d.active = s.active && !s.parked && (s.phase32 != 0) == fp32;
      if (!d.active) return;
      auto slotU = [&](int q) { return (void*)((char*)s.U.p + (size_t)q * mP * sz); };
      const size_t szr = fp32 && pb.r8 ? 1 : (fp32 && pb.r16 ? 2 : sz);  // FP8 / BF16 history
      auto slotR = [&](int q) { return (void*)((char*)s.R.p + (size_t)q * mP * szr); };
      d.t = fp32 ? (void*)s.dcur : (void*)s.t64.p;
      d.ncomb = (int)s.coef.size();
      const bool u16 = fp32 && pb.u16;
      auto slotD = [&](int q) { return (void*)((uint16_t*)s.U.p + (size_t)q * mP); };
      if (u16) {
        // t = sum_j c_j u_j = u_newest - sum_i (c_0 + .. + c_i) (u_i+1 - u_i),
        // the difference u_i+1 - u_i in the BF16 slot of entry i + 1
        d.u16 = 1;
        d.u[0] = s.ucur;
        d.coef[0] = 1.0;
        double cum = 0.0;
        for (int i = 0; i + 1 < d.ncomb; ++i) {
          cum += s.coef[i];
          d.u[1 + i] = slotD(s.order[i + 1]);
          d.coef[1 + i] = -cum;
        }
      } else {
        for (int j = 0; j < d.ncomb; ++j) {
          d.u[j] = slotU(s.order[j]);
          d.coef[j] = s.coef[j];
          if (fp32 && s.sbase[s.order[j]] != s.base) d.shift_w += s.coef[j];
        }
      }
      d.shift = s.dshift;
      // slot for the new r, u: a free one, else the oldest (dropped)
      s.dotted = s.order;
      if ((int)s.order.size() < Dp) {
        int q = 0;
        while (std::find(s.order.begin(), s.order.end(), q) != s.order.end()) ++q;
        s.slot_new = q;
      } else {
        s.slot_new = s.order.front();
        s.dotted.erase(s.dotted.begin());
      }
      d.ndot = (int)s.dotted.size();
      for (int j = 0; j < d.ndot; ++j) d.rh[j] = slotR(s.dotted[j]);
      if (fp32 && pb.r8) {
        d.r8 = 1;
        d.rsc = fp8_scale(s.resid);
        s.rsc[s.slot_new] = d.rsc;
        for (int j = 0; j < d.ndot; ++j) d.rhinv[j] = 1.0f / s.rsc[s.dotted[j]];
      }
      d.r = slotR(s.slot_new);
      d.unew = u16 ? (void*)s.ucur : slotU(s.slot_new);
      d.dnew = u16 ? slotD(s.slot_new) : nullptr;
      d.lam = fp32 ? lam32(pb, cf) : cf.lam;
      d.bv = fp32 ? (const void*)s.bv32.p : (const void*)s.bv64.p;
      d.W = s.W.p;
      d.Fs = fp32 ? (const void*)s.fs32.p : (const void*)s.fs64.p;
      d.part = s.part.p;
      // FP32 passes iterate the correction delta around the FP64 base:
      // mode 5 (closure difference), or for the linear adjoint mode 1 with
      // the base's FP64 residual as its constant term: D K delta + r_b - delta
      d.mode = fp32 ? (cf.mode == 1 ? 1 : 5) : cf.mode;
      d.dt = s.Dt.p;
      d.gv = (fp32 && cf.mode == 1) ? s.r64.p : s.gmu.p;
      d.fb = s.fb32.p;
      d.r16 = fp32 && pb.r16;
    };
    if (fp32 && dev_ok) {
      // The MDIIS step on the device: the solutes' history state goes up
      // once; after that the host takes back only the solutes that ask for
      // a rebase or fail, and sends them up again after it.
      if (!dev_on) {
        if (!pb.d_mdiis.p) {
          pb.d_sd32 = DevBuf<SoluteDesc>(B);
          pb.d_mdiis = DevBuf<MdiisDev>(B);
          HIPCHK(hipHostMalloc((void**)&pb.h_status.p, B * sizeof(MdiisStatus),
                               hipHostMallocCoherent));
        }
        std::memset(pb.h_status.p, 0, B * sizeof(MdiisStatus));
        dev_dirty.clear();
        for (int b = 0; b < B; ++b) dev_dirty.push_back(b);
        dev_on = true;
      }
      if (!dev_dirty.empty()) {
        for (int b : dev_dirty) {
          make_desc(b);
          if (sd[b].active) pack(b, Dp);
          next_hist[b] = std::max(sd[b].ncomb, sd[b].ndot);
        }
        load_rows(dev_dirty);
        dev_dirty.clear();
      }
      int hist = 0;
      for (int b = 0; b < B; ++b)
        if (pb.st[b].active && pb.st[b].phase32 && !pb.st[b].parked)
          hist = std::max(hist, next_hist[b]);
      evaluate(pb, true, sd, 1, hist);
      wait_status();
      ++npass32;
      std::vector<int> rebase_mask(B, 0), back;
      bool any_rebase = false;
      for (int b = 0; b < B; ++b) {
        auto& s = pb.st[b];
        if (!(s.active && s.phase32) || s.parked) continue;
        const MdiisStatus& q = pb.h_status.p[b];
        s.iters = q.iters;
        s.iters32 = q.iters32;
        s.resid = q.resid;
        next_hist[b] = q.hist;
        if (cf.verbose)
          std::fprintf(stderr, "rism3d: solute %d iter %d fp32 |F|inf = %.3e |F|2 = %.3e\n", b,
                       s.iters, q.resid, std::sqrt(q.r2));
        if (q.flag == 2) {
          s.status = RISM3D_NAN;
          s.active = 0;
          back.push_back(b);
        } else if (q.flag == 1) {
          rebase_mask[b] = 1;
          any_rebase = true;
          back.push_back(b);
        } else if (q.flag == 4 || q.flag == 5) {
          if (s.ready_at < 0) s.ready_at = npass32;
          s.parked = q.flag == 5;
        }
      }
      group_rebases(rebase_mask, any_rebase, &back);
      if (!back.empty()) {  // their state back to the host
        const size_t bytes = (size_t)B * sizeof(MdiisDev);
        int k = 0;
        void* h = pb.down.slot(bytes, &k);
        HIPCHK(hipMemcpyAsync(h, pb.d_mdiis.p, bytes, hipMemcpyDeviceToHost, pb.s));
        pb.down.mark(k, pb.s);
        HIPCHK(hipStreamSynchronize(pb.s));
        for (int b : back) unpack(b, static_cast<const MdiisDev*>(h)[b]);
      }
      if (any_rebase) {
        std::vector<double> rinf, r2;
        rebase(pb, cf, rebase_mask, true, rinf, r2);
        after_rebase(rebase_mask, rinf, r2, false);
        for (int b = 0; b < B; ++b)
          if (rebase_mask[b]) dev_dirty.push_back(b);
      }
      continue;
    }
    for (int b = 0; b < B; ++b) make_desc(b);
    evaluate(pb, fp32, sd);
    download_out(pb, out);
    HIPCHK(hipGetLastError());

    std::vector<int> rebase_mask(B, 0);
    bool any_rebase = false;
    for (int b = 0; b < B; ++b) {
      auto& s = pb.st[b];
      if (!sd[b].active) continue;
      const double* o = out.data() + (size_t)b * NPART;
      const double rinf = o[0], r2 = o[1];
      s.iters++;
      if (fp32) s.iters32++;
      s.resid = rinf;
      if (cf.verbose)
        std::fprintf(stderr, "rism3d: solute %d iter %d %s |F|inf = %.3e |F|2 = %.3e\n", b, s.iters,
                     fp32 ? "fp32" : "fp64", rinf, std::sqrt(r2));
      if (!std::isfinite(rinf) || !std::isfinite(r2)) {
        s.status = RISM3D_NAN;
        s.active = 0;
        continue;
      }
      if (fp32) {
        // Rebase once the correction's residual fell 1e3 below F(t_b), or
        // below mixed_tol / 2 (the FP64 test then decides), or stopped
        // falling (no 30% gain on its best in 12 passes, counted below
        // 1e-2 F(t_b)), or at the iteration limit.
        // FP32 only: no rebase until the floor (stall below 1e-3) or
        // mixed_tol / 2.
        const bool only32 = cf.precision == RISM3D_FP32;
        if (rinf < 0.7 * s.best32) { s.best32 = rinf; s.stall32 = 0; }
        else if (rinf <= (only32 ? 1e-3 : 1e-2 * s.rb_rinf)) { s.stall32++; }
        // The new entry still enters the history (the rebase moves it).
        // ready (with a rebase wait): the 1e3 drop, or mixed_tol / 2 reached
        const bool drop = !only32 && rinf <= 1e-3 * s.rb_rinf, conv = rinf <= 0.5 * s.tol_mixed;
        if (s.stall32 >= 12 || s.iters32 >= cf.max_iter || (pb.rebase_wait <= 0 && (drop || conv)) ||
            (only32 && conv)) {
          rebase_mask[b] = 1;
          any_rebase = true;
        } else if (drop || conv) {
          if (s.ready_at < 0) s.ready_at = npass32 + 1;
          s.parked = conv;
        }
        s.sbase[s.slot_new] = s.base;
      } else {
        if (cf.mode == 1 && s.iters == 1) s.tol_lin = cf.rel_tol * rinf;  // r(0) = g
        // FP64 and the MIXED polish both aim at norm_tol; the polish stops
        // after polish_iters steps and is OK if it reached mixed_tol.
        const double tol = cf.mode == 1 ? s.tol_lin : cf.tol;
        if (rinf <= tol) {
          s.status = RISM3D_OK;
          s.active = 0;
          continue;
        }
        if (cf.precision == RISM3D_MIXED && s.polish >= cf.polish) {
          // no polish: the FP32 phase's own end decides (mixed_tol or its
          // floor reached); a polish must reach mixed_tol in FP64
          const bool ok = cf.polish == 0 ? s.reached_mixed != 0 : rinf <= cf.mixed_tol;
          s.status = ok ? RISM3D_OK : RISM3D_NONCONVERGED;
          s.active = 0;
          continue;
        }
        s.polish++;
        if (s.iters - s.iters32 >= cf.max_iter) {
          s.status = RISM3D_NONCONVERGED;
          s.active = 0;
          continue;
        }
      }
      // enter the new entry, refresh its Gram row
      s.order = s.dotted;
      s.order.push_back(s.slot_new);
      for (int j = 0; j < (int)s.dotted.size(); ++j) {
        s.gram[s.slot_new * D + s.dotted[j]] = o[2 + j];
        s.gram[s.dotted[j] * D + s.slot_new] = o[2 + j];
      }
      s.gram[s.slot_new * D + s.slot_new] = r2;
      // restart when the residual blows up against the best one, or stalls
      if (r2 < s.best) { s.best = r2; s.since_best = 0; } else { s.since_best++; }
      if (r2 > 100.0 * s.best || s.since_best > 3 * Dp) {
        s.order.assign(1, s.slot_new);
        s.best = r2;
        s.since_best = 0;
      }
      const int nh = (int)s.order.size();
      std::vector<double> G(nh * nh);
      for (int i = 0; i < nh; ++i)
        for (int j = 0; j < nh; ++j) G[i * nh + j] = s.gram[s.order[i] * D + s.order[j]];
      if (nh == 1 || !diis_coefficients(G, nh, s.coef)) {
        s.order.assign(1, s.slot_new);
        s.coef.assign(1, 1.0);
      }
    }
    if (fp32) {
      ++npass32;
      group_rebases(rebase_mask, any_rebase, nullptr);
    }
    if (any_rebase) {
      std::vector<double> rinf, r2;
      rebase(pb, cf, rebase_mask, true, rinf, r2);
      after_rebase(rebase_mask, rinf, r2, false);
    }
  }
  // solutes that stopped inside the FP32 phase (NaN): t = t_b + delta
  for (int b = 0; b < B; ++b)
    if (pb.st[b].phase32) {
      const int ch = (int)std::min<int64_t>(2048, (mP + NT - 1) / NT);
      hipLaunchKernelGGL(add_delta, dim3(ch), dim3(NT), 0, pb.s, pb.st[b].t64.p, pb.st[b].dcur, mP);
      pb.st[b].phase32 = 0;
    }
  HIPCHK(hipStreamSynchronize(pb.s));
}

// ------------------------------------------------ the reference's solvers
// Newton (bgy3d_snes_newton, bgy3d-snes.c:161-346), Jager
// (bgy3d_snes_jager, :402-519) and trial (:387-399), FP64, one solute at a
// time (the others inactive).  Each keeps its state in t64 of the solute and
// counts every full pass (residual or Jacobian application) in iters.
struct OneSolute {
  Problem& pb;
  int b;
  std::vector<SoluteDesc> sd;
  std::vector<double> out;
  OneSolute(Problem& p, int bb) : pb(p), b(bb), sd(p.B), out((size_t)p.B * NPART) {}
  SoluteDesc& desc() {
    for (auto& d : sd) std::memset(&d, 0, sizeof d);
    auto& s = pb.st[b];
    SoluteDesc& d = sd[b];
    d.active = 1;
    d.bv = s.bv64.p;
    d.W = s.W.p;
    d.Fs = s.fs64.p;
    d.part = s.part.p;
    return d;
  }
  // evaluate the prepared descriptor; returns (max|r|, ||r||^2)
  std::pair<double, double> run() {
    evaluate(pb, false, sd);
    download_out(pb, out);
    HIPCHK(hipGetLastError());
    pb.st[b].iters++;
    return {out[(size_t)b * NPART], out[(size_t)b * NPART + 1]};
  }
  // F at t = u0 + a u1 (ncomb 2) or at t itself (u1 == nullptr) into r;
  // t64 receives the point
  std::pair<double, double> F(const double* u0, const double* u1, double a, double* r) {
    SoluteDesc& d = desc();
    d.t = pb.st[b].t64.p;
    if (u1) {
      d.ncomb = 2;
      d.u[0] = u0;
      d.u[1] = u1;
      d.coef[0] = 1.0;
      d.coef[1] = a;
    }
    d.r = r;
    return run();
  }
};

// Picard x += lam F(x) with the step schedule of bgy3d_snes_jager when
// jager, else fixed lam; at most nsteps residuals.  Returns true when
// max|F| <= tol.
bool picard_steps(Problem& pb, const Config& cf, int b, double lam, int nsteps, bool jager) {
  OneSolute os(pb, b);
  auto& s = pb.st[b];
  DevBuf<double> r(pb.mP);
  double a0 = lam, a1 = lam, a_prev = 0.0, norm_old = 0.0;
  int mycount = 0, upwards = 0;
  const int nth = 10;
  for (int iter = 0; iter < nsteps; ++iter) {
    auto [rinf, r2] = os.F(s.t64.p, iter ? r.p : nullptr, a_prev, r.p);
    s.resid = rinf;
    if (cf.verbose)
      std::fprintf(stderr, "rism3d: solute %d iter %d %s |F|inf = %.3e step %.3g\n", b, s.iters,
                   jager ? "jager" : "picard", rinf, a_prev);
    if (!std::isfinite(rinf) || !std::isfinite(r2)) { s.status = RISM3D_NAN; return false; }
    if (rinf <= cf.tol) { s.status = RISM3D_OK; return true; }
    if (!jager) { a_prev = lam; continue; }
    const double a = (iter % nth == 0) ? a1 : a0;
    const bool up = rinf > norm_old;
    norm_old = rinf;
    mycount++;
    if (iter % nth != 1) upwards = up ? 1 : 0;
    if (iter % nth == 1 && up && upwards == 0 && iter > 2 * nth) {
      a1 = std::max(a1 / 2.0, a0);
      mycount = 0;
    }
    if (mycount > 2 * nth) {
      a1 = std::min(a1 * 2.0, 1.0);
      mycount = 0;
    }
    a_prev = a;
  }
  // the last step is applied, as the reference's loop does
  if (nsteps > 0 && a_prev != 0.0) {
    const int ch = (int)std::min<int64_t>(2048, (pb.mP + NT - 1) / NT);
    hipLaunchKernelGGL(vcomb, dim3(ch), dim3(NT), 0, pb.s, s.t64.p, r.p, a_prev, 1.0, pb.mP);
    // vcomb computes y = a x + b y: t64 = a_prev r + t64
    HIPCHK(hipStreamSynchronize(pb.s));
  }
  s.status = RISM3D_NONCONVERGED;
  return false;
}

// Newton with a matrix-free Jacobian J v = K (D v) - v (modes 6 and the
// dc/dt field D), GMRES(30) with classical Gram-Schmidt applied twice, the
// linear solve to 1e-3 relative (the reference's KSP rtol, 100 x PETSc's
// 1e-5, bgy3d-snes.c:274), and a backtracking line search on ||F||_2
// (SNESNEWTONLS: sufficient decrease 1e-4, halving, 10 tries).
void newton_solve(Problem& pb, const Config& cf, int b, int max_newton) {
  const int R = 30;
  const double ksp_rtol = 1e-3;
  auto& s = pb.st[b];
  const int64_t n = pb.mP;
  const int ch = (int)std::min<int64_t>(1024, (n + NT - 1) / NT);
  OneSolute os(pb, b);
  DevBuf<double> V((size_t)(R + 1) * n), w(n), r(n), t0(n), dx(n), D(n), part((size_t)ch * KMAX),
      hdev(KMAX);
  std::vector<double> hp((size_t)ch * KMAX);
  auto dots = [&](const double* Vb, int nv, const double* wv, std::vector<double>& res) {
    hipLaunchKernelGGL(mdot, dim3(ch), dim3(NT), 0, pb.s, Vb, nv, wv, n, part.p);
    HIPCHK(hipMemcpyAsync(hp.data(), part.p, (size_t)ch * nv * 8, hipMemcpyDeviceToHost, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
    res.assign(nv, 0.0);
    for (int c = 0; c < ch; ++c)
      for (int i = 0; i < nv; ++i) res[i] += hp[(size_t)c * nv + i];
  };
  auto axpy_many = [&](double* wv, const double* Vb, const std::vector<double>& h, int nv) {
    HIPCHK(hipMemcpy(hdev.p, h.data(), nv * 8, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(maxpy, dim3(ch), dim3(NT), 0, pb.s, wv, Vb, hdev.p, nv, n);
  };
  auto [rinf, r2] = os.F(s.t64.p, nullptr, 0.0, r.p);
  for (int it = 0;; ++it) {
    s.resid = rinf;
    if (cf.verbose)
      std::fprintf(stderr, "rism3d: solute %d newton %d |F|inf = %.3e |F|2 = %.3e\n", b, it, rinf,
                   std::sqrt(r2));
    if (!std::isfinite(rinf) || !std::isfinite(r2)) { s.status = RISM3D_NAN; return; }
    if (rinf <= cf.tol) { s.status = RISM3D_OK; return; }
    if (it >= max_newton || s.iters >= cf.max_iter * 40) { s.status = RISM3D_NONCONVERGED; return; }
    // D = dc/dt at t; GMRES on J dx = -r from dx = 0
    hipLaunchKernelGGL(closure_dcdt, dim3(ch), dim3(NT), 0, pb.s, pb.g, s.t64.p, s.bv64.p, D.p, n);
    const double beta = std::sqrt(r2);
    hipLaunchKernelGGL(vcomb, dim3(ch), dim3(NT), 0, pb.s, V.p, r.p, -1.0 / beta, 0.0, n);
    std::vector<std::vector<double>> H(R + 1, std::vector<double>(R, 0.0));
    std::vector<double> cs(R, 0.0), sn(R, 0.0), gg(R + 1, 0.0);
    gg[0] = beta;
    int k = 0;
    for (int j = 0; j < R; ++j) {
      SoluteDesc& d = os.desc();
      d.t = V.p + (size_t)j * n;
      d.r = w.p;
      d.dt = D.p;
      d.mode = 6;
      os.run();
      std::vector<double> h, h2;
      dots(V.p, j + 1, w.p, h);
      axpy_many(w.p, V.p, h, j + 1);
      dots(V.p, j + 1, w.p, h2);
      axpy_many(w.p, V.p, h2, j + 1);
      for (int i = 0; i <= j; ++i) H[i][j] = h[i] + h2[i];
      std::vector<double> nn;
      dots(w.p, 1, w.p, nn);
      const double hn = std::sqrt(std::max(nn[0], 0.0));
      H[j + 1][j] = hn;
      if (hn > 0)
        hipLaunchKernelGGL(vcomb, dim3(ch), dim3(NT), 0, pb.s, V.p + (size_t)(j + 1) * n, w.p,
                           1.0 / hn, 0.0, n);
      for (int i = 0; i < j; ++i) {  // earlier rotations
        const double a = H[i][j], c = H[i + 1][j];
        H[i][j] = cs[i] * a + sn[i] * c;
        H[i + 1][j] = -sn[i] * a + cs[i] * c;
      }
      const double den = std::hypot(H[j][j], H[j + 1][j]);
      cs[j] = den > 0 ? H[j][j] / den : 1.0;
      sn[j] = den > 0 ? H[j + 1][j] / den : 0.0;
      H[j][j] = den;
      H[j + 1][j] = 0.0;
      gg[j + 1] = -sn[j] * gg[j];
      gg[j] = cs[j] * gg[j];
      k = j + 1;
      if (std::fabs(gg[j + 1]) <= ksp_rtol * beta || hn == 0.0) break;
    }
    std::vector<double> y(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double acc = gg[i];
      for (int q = i + 1; q < k; ++q) acc -= H[i][q] * y[q];
      y[i] = H[i][i] != 0.0 ? acc / H[i][i] : 0.0;
    }
    // dx = sum_i y_i V_i
    hipLaunchKernelGGL(vcomb, dim3(ch), dim3(NT), 0, pb.s, dx.p, V.p, y[0], 0.0, n);
    if (k > 1) {
      std::vector<double> my(k - 1);
      for (int i = 1; i < k; ++i) my[i - 1] = -y[i];
      axpy_many(dx.p, V.p + n, my, k - 1);
    }
    // line search from t0 = t
    HIPCHK(hipMemcpyAsync(t0.p, s.t64.p, n * 8, hipMemcpyDeviceToDevice, pb.s));
    double alpha = 1.0;
    bool accepted = false;
    for (int ls = 0; ls < 10; ++ls) {
      auto [ri, rr] = os.F(t0.p, dx.p, alpha, r.p);
      if (std::isfinite(ri) && std::isfinite(rr) && rr <= (1.0 - 2e-4 * alpha) * r2) {
        rinf = ri;
        r2 = rr;
        accepted = true;
        break;
      }
      alpha *= 0.5;
    }
    if (!accepted) {
      HIPCHK(hipMemcpyAsync(s.t64.p, t0.p, n * 8, hipMemcpyDeviceToDevice, pb.s));
      HIPCHK(hipStreamSynchronize(pb.s));
      s.status = RISM3D_NONCONVERGED;  // line search failed (SNES would diverge)
      return;
    }
  }
}

void reference_solver(Problem& pb, const Config& cf) {
  for (int b = 0; b < pb.B; ++b) {
    auto& s = pb.st[b];
    if (!s.active) continue;
    if (cf.solver == RISM3D_SOLVER_JAGER) {
      picard_steps(pb, cf, b, cf.lam, cf.max_iter, true);
    } else {
      bool done = false;
      if (cf.solver == RISM3D_SOLVER_TRIAL) {
        Config c = cf;
        done = picard_steps(pb, c, b, 0.02, 100, false);
      }
      if (!done && s.status != RISM3D_NAN) newton_solve(pb, cf, b, cf.max_iter);
    }
    s.active = 0;
  }
}

// Closure continuation for solutes whose iteration diverged (max|F| not
// finite): HNC and PSE-n (n >= 2) grow like exp(x) or x^n where the solute
// field is strongly attractive, and from t = 0 the first steps can
// overshoot into overflow.  Such solutes restart from t = 0 along
// KH = PSE-1, PSE-2, ..., each rung converged loosely (max|F| <= 1e-2) and
// handed to the next, the target closure last, to the full tolerance.
// For HNC the ladder ends with PSE-4.  The fixed point is the target
// closure's: only the starting point of its iteration changes.
// Iteration counts add up over the rungs.
void continuation(Problem& pb, const Config& cf, int target) {
  std::vector<std::pair<int, int>> rungs;  // (closure enum, PSE order)
  const int top = target == RISM3D_CLOSURE_HNC ? 5 : target - RISM3D_CLOSURE_PSE0;
  if (target != RISM3D_CLOSURE_HNC && target < RISM3D_CLOSURE_PSE2) return;
  rungs.push_back({RISM3D_CLOSURE_KH, 1});
  for (int n = 2; n < top; ++n) rungs.push_back({RISM3D_CLOSURE_PSE0 + n, n});
  rungs.push_back({target, target == RISM3D_CLOSURE_HNC ? 0 : top});
  const int B = pb.B;
  std::vector<int> redo(B, 0), tot(B, 0), tot32(B, 0);
  bool any = false;
  for (int b = 0; b < B; ++b) {
    auto& s = pb.st[b];
    s.active = 0;
    if (s.status != RISM3D_NAN) continue;
    redo[b] = 1;
    any = true;
    tot[b] = s.iters;
    tot32[b] = s.iters32;
    HIPCHK(hipMemsetAsync(s.t64.p, 0, pb.mP * 8, pb.s));
  }
  if (!any) return;
  const std::vector<int> joined = redo;
  const GridArgs g0 = pb.g;
  for (size_t k = 0; k < rungs.size(); ++k) {
    const bool last = k + 1 == rungs.size();
    pb.g.closure = rungs[k].first;
    pb.g.pse_order = rungs[k].second;
    for (int b = 0; b < B; ++b) {
      auto& s = pb.st[b];
      if (!redo[b]) continue;
      s.active = 1;
      s.status = RISM3D_NONCONVERGED;
      s.iters = s.iters32 = 0;
      s.reached_mixed = 0;
      s.best32 = INFINITY;
      s.stall32 = 0;
      s.polish = 0;
      s.resid = s.best = INFINITY;
      s.since_best = 0;
    }
    Config c = cf;
    if (!last) {
      c.tol = std::max(cf.tol, 1e-2);
      c.mixed_tol = std::max(cf.mixed_tol, 1e-2);
      c.polish = 0;
    }
    if (cf.verbose)
      std::fprintf(stderr, "rism3d: continuation rung %d, closure %d\n", (int)k, rungs[k].first);
    iterate_batch(pb, c);
    for (int b = 0; b < B; ++b) {
      if (!redo[b]) continue;
      auto& s = pb.st[b];
      tot[b] += s.iters;
      tot32[b] += s.iters32;
      if (s.status == RISM3D_NAN) redo[b] = 0;  // diverged on a rung: stays NAN
    }
  }
  pb.g = g0;
  for (int b = 0; b < B; ++b)
    if (joined[b]) {
      pb.st[b].iters = tot[b];
      pb.st[b].iters32 = tot32[b];
    }
}

struct Inputs {
  const rism3d_problem* pd;
  int m;
  const rism3d_site* solvent;
  int B;
  const int* n_sites;
  const rism3d_site* const* solutes;
  const double* chi_k;
  rism3d_options opt;
  const double* t_eval = nullptr;  // residual_only: t in
  double* r_out = nullptr;         // residual_only: F(t) out
  const double* j_dt = nullptr;    // jacobian: dt in (t in t_eval), J dt out in r_out
  int j_flag = 0;
  int noalloc = 0;                 // every device buffer from the installed arena
  hipStream_t stream = nullptr;    // a stream the caller keeps (else one per call)
};

int validate(const Inputs& in) {
  const rism3d_problem* pd = in.pd;
  if (!pd || !in.solvent || in.m < 1 || in.m > MAX_SITES || in.B < 1) return RISM3D_BAD_ARG;
  for (int d = 0; d < 3; ++d) {
    if (!supported_n(pd->N[d])) return RISM3D_UNSUPPORTED;
    if (!(pd->L[d] > 0)) return RISM3D_BAD_ARG;
  }
  if (!(pd->rmax > 0) || pd->nrad < 4 || !(pd->beta > 0) || !(pd->rho > 0)) return RISM3D_BAD_ARG;
  if (pd->closure < 0 || pd->closure > RISM3D_CLOSURE_PSE7) return RISM3D_BAD_ARG;
  for (int b = 0; b < in.B; ++b)
    if (in.n_sites[b] < 1 || !in.solutes[b]) return RISM3D_BAD_ARG;
  // the k table must reach the grid corner: interp asserts otherwise
  // (bgy3d-vec.c:31-32)
  double k2 = 0;
  for (int d = 0; d < 3; ++d) {
    const double kd = (pd->N[d] / 2) * 2 * M_PI / pd->L[d];
    k2 += kd * kd;
  }
  const double dk = M_PI / pd->rmax;
  const int imax = (int)(std::sqrt(k2) / dk - 0.5 + 0.5);
  if (!in.opt.chi3d && imax > pd->nrad - 1) return RISM3D_BAD_ARG;
  if (in.opt.diis_depth < 1 || in.opt.diis_depth > MAX_DIIS) return RISM3D_BAD_ARG;
  if (in.opt.solver < RISM3D_SOLVER_MDIIS || in.opt.solver > RISM3D_SOLVER_TRIAL)
    return RISM3D_BAD_ARG;
  if (in.opt.solver >= RISM3D_SOLVER_NEWTON && in.opt.precision != RISM3D_FP64)
    return RISM3D_UNSUPPORTED;
  if ((in.opt.solver == RISM3D_SOLVER_NEWTON || in.opt.solver == RISM3D_SOLVER_TRIAL) &&
      pd->closure == RISM3D_CLOSURE_PY)
    return RISM3D_UNSUPPORTED;  // no closure differential for PY (closures.f90:131-153)
  // z-pass LDS: m padded lines per kx column, at least 2 columns
  {
    const size_t lds64 =
        (size_t)in.m * zstar_lines<double>(pd->N[2], in.m) * lstride_for(pd->N[2]) * 16;
    if (lds64 > 64 * 1024) return RISM3D_UNSUPPORTED;
  }
  return RISM3D_OK;
}


// Gradients after a converged solve (t in t64 of every solute).
// TPT ('free-energy-gradient, hnc3d.c:1666-1742) for res[b].gradient, the
// exact adjoint d mu / d(x, y, z, q, sigma, epsilon) for res[b].dmu_dsite,
// and the reference's linear-response gradient ('free-energy-response,
// hnc3d.c:1969-2134) for res[b].response.  The last two share one linear
// solve: lambda = D_t K lambda + dmu/dt by linear MDIIS, then K lambda.
// The reference's response path solves J dt = beta J' dv once per site
// and axis (3 n solves); by J^T = D_t K - 1 (K symmetric) its result is
// sum_i <A_i | dv_i> with A_i = -beta (D_i + 1) (K lambda)_i + dmu/dv_i,
// one field reduction.  The adjoint adds Lambda = r2c(lambda) and the
// tau and u_c terms (grad_tau, grad_field mode 1).
// info / print_info (bgy3d-potential.c:341-491, cage = false) for solute b
// after its solve: rho_v, its Poisson potential U_v (a pass of the residual
// pipeline with one "site", the closure step copying rho_v and the k-space
// step the Poisson factor), the grid sums and the interpolated U_v at the
// sites.  W and R slot 0 of the solute are reused as work space.
void info_observables(Problem& pb, const Inputs& in, int b, const SitesDesc& sites,
                      rism3d_info* out) {
  GridArgs& g = pb.g;
  const rism3d_problem& pd = *in.pd;
  auto& s = pb.st[b];
  const int n = in.n_sites[b];
  const rism3d_site* su = in.solutes[b];
  const int64_t P = g.P;
  DevBuf<double> rhov(P), ve(P);
  const int cp = (int)std::min<int64_t>(2048, (P + NT - 1) / NT);
  hipLaunchKernelGGL(info_rho_v, dim3(cp), dim3(NT), 0, pb.s, g, s.t64.p, s.bv64.p, pd.rho,
                     rhov.p);
  // U_v = IFFT(FFT(rho_v) f) / L^3, f = 4 pi eps0^-1 h^3 / k^2
  std::vector<SoluteDesc> sd(pb.B);
  std::memset(sd.data(), 0, sd.size() * sizeof(SoluteDesc));
  sd[b].active = 1;
  sd[b].t = rhov.p;
  sd[b].bv = s.bv64.p;
  sd[b].W = s.W.p;
  sd[b].part = s.part.p;
  sd[b].r = ve.p;
  sd[b].mode = 2;
  const int m_save = g.m, special_save = g.special;
  g.m = 1;
  g.special = 2;
  evaluate(pb, false, sd);
  // Y = FFT(rho_v) alone in W for the interpolation (forward passes)
  g.special = 0;
  sd[b].mode = 3;
  HIPCHK(hipMemcpyAsync(pb.d_sd.p, sd.data(), sd.size() * sizeof(SoluteDesc),
                        hipMemcpyHostToDevice, pb.s));
  {
    using T = double;
    constexpr int LN = y_lines<T>();
    const dim3 gx(resid_grid_x(g), 1, pb.B);
    const dim3 gy((g.NX + LN - 1) / LN, g.N2, pb.B);
    launch_closure_r2c<T>(g.N0, gx, pb.s, pb.d_sd.p, g, pb.tw.d[1].p, pb.tw.d[0].p, hist_of(sd));
    launch_fft_y<T, false>(g.N1, gy, pb.s, pb.d_sd.p, g, pb.tw.d[2].p);
    launch_zstar<T>(g.N2, g.NX, g.N1, pb.B, pb.s, pb.d_sd.p, g, pb.tw.d[3].p, pb.chi3.p,
                    pb.tau3.p);
  }
  g.m = m_save;
  g.special = special_save;
  DevBuf<double> part((size_t)cp * INFO_NSUM);
  hipLaunchKernelGGL(info_sums, dim3(cp), dim3(NT), 0, pb.s, g, sites, rhov.p, ve.p, pd.L[0],
                     pd.L[1], pd.L[2], pb.h[0], pb.h[1], pb.h[2], part.p);
  std::vector<double> xs(3 * (size_t)n);
  for (int a = 0; a < n; ++a)
    for (int d = 0; d < 3; ++d) xs[3 * a + d] = su[a].x[d];
  DevBuf<double> d_xs(xs.size());
  HIPCHK(hipMemcpyAsync(d_xs.p, xs.data(), xs.size() * 8, hipMemcpyHostToDevice, pb.s));
  const int ck = (int)std::min<int64_t>(512, (g.K + NT - 1) / NT);
  DevBuf<double> ipart((size_t)n * ck);
  hipLaunchKernelGGL(info_interp, dim3(ck, n), dim3(NT), 0, pb.s, g, (const cpx<double>*)s.W.p,
                     d_xs.p, pd.L[0], pd.L[1], pd.L[2], pb.h[0], pb.h[1], pb.h[2], ipart.p);
  std::vector<double> hp((size_t)cp * INFO_NSUM), hi((size_t)n * ck);
  HIPCHK(hipMemcpyAsync(hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost, pb.s));
  HIPCHK(hipMemcpyAsync(hi.data(), ipart.p, hi.size() * 8, hipMemcpyDeviceToHost, pb.s));
  if (out->potential)
    HIPCHK(hipMemcpyAsync(out->potential, ve.p, P * 8, hipMemcpyDeviceToHost, pb.s));
  HIPCHK(hipStreamSynchronize(pb.s));
  HIPCHK(hipGetLastError());
  double acc[INFO_NSUM];
  for (int q = 0; q < INFO_NSUM; ++q) {
    double sum = 0.0, comp = 0.0;
    for (int c = 0; c < cp; ++c) {
      const double v = hp[(size_t)c * INFO_NSUM + q];
      const double t = sum + v;
      comp += (std::fabs(sum) >= std::fabs(v)) ? (sum - t) + v : (v - t) + sum;
      sum = t;
    }
    acc[q] = sum + comp;
  }
  const double* h = pb.h;
  const double dV = h[0] * h[1] * h[2];
  auto moments = [&](const double* v, double* q, double* d, double (*Q)[3]) {
    *q = dV * v[0];
    for (int a = 0; a < 3; ++a) d[a] = dV * h[a] * v[1 + a];
    const int idx[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) Q[a][c] = dV * h[a] * h[c] * v[idx[a][c]];
  };
  moments(acc, &out->q_v, out->d_v, out->Q_v);
  moments(acc + 10, &out->q_u, out->d_u, out->Q_u);
  out->U_v_rho_u = dV * acc[20];
  out->rho_v_U_u = dV * acc[21];
  for (int d = 0; d < 3; ++d) {
    double dn = 0.0;
    for (int a = 0; a < n; ++a) dn += su[a].charge * su[a].x[d];
    out->d_n[d] = dn;
  }
  double val1 = 0.0;
  for (int a = 0; a < n; ++a) {
    double sum = 0.0;
    for (int c = 0; c < ck; ++c) sum += hi[(size_t)a * ck + c];
    if (out->v_sites) out->v_sites[a] = sum;
    val1 += su[a].charge * sum;
  }
  out->U_v_rho_N = val1;
}

void gradients(Problem& pb, const Inputs& in, rism3d_result* res, const Config& cf0,
               const std::vector<SitesDesc>& sites) {
  const GridArgs& g = pb.g;
  const rism3d_problem& pd = *in.pd;
  const int B = pb.B;
  const int64_t mP = pb.mP;
  std::vector<double> sse(3 * g.m);
  for (int i = 0; i < g.m; ++i) {
    sse[3 * i] = in.solvent[i].sigma;
    sse[3 * i + 1] = in.solvent[i].epsilon;
    sse[3 * i + 2] = in.solvent[i].charge;
  }
  bool want_tpt = false, want_lam = false;
  // the reference's closure1 has no PY branch (closures.f90:131-153), so
  // it has no response for PY; the adjoint covers the tau and Ng paths
  auto wants_adj = [&](int b) { return res[b].dmu_dsite != nullptr; };
  auto wants_resp = [&](int b) { return res[b].response != nullptr && g.closure != 2; };
  for (int b = 0; b < B; ++b) {
    const int n = sites[b].n;
    if (res[b].dmu_dsite && !wants_adj(b))
      for (int q = 0; q < n * 6; ++q) res[b].dmu_dsite[q] = NAN;
    if (res[b].response && !wants_resp(b))
      for (int q = 0; q < n * 3; ++q) res[b].response[q] = NAN;
    if (pb.st[b].status != RISM3D_OK) continue;
    want_tpt |= res[b].gradient != nullptr;
    want_lam |= wants_adj(b) || wants_resp(b);
  }
  if (!want_tpt && !want_lam) return;
  DevBuf<double> d_sse(sse.size());
  HIPCHK(hipMemcpyAsync(d_sse.p, sse.data(), sse.size() * 8, hipMemcpyHostToDevice, pb.s));
  const int cp = (int)std::min<int64_t>(512, (g.P + NT - 1) / NT);
  const int ck = (int)std::min<int64_t>(512, (g.K + NT - 1) / NT);
  DevBuf<double> A(mP), Bf(g.P);

  auto field_reduce = [&](int b, int mode, std::vector<double>& acc6) {
    const int n = sites[b].n;
    DevBuf<double> part((size_t)n * cp * 6);
    hipLaunchKernelGGL(grad_field, dim3(cp, n), dim3(NT), 0, pb.s, g, sites[b], d_sse.p, A.p, Bf.p,
                       pd.L[0], pd.L[1], pd.L[2], pb.h[0], pb.h[1], pb.h[2], mode, part.p);
    std::vector<double> hp(part.n);
    HIPCHK(hipMemcpyAsync(hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
    acc6.assign((size_t)n * 6, 0.0);
    for (int a = 0; a < n; ++a)
      for (int c = 0; c < cp; ++c)
        for (int p = 0; p < 6; ++p) acc6[a * 6 + p] += hp[((size_t)a * cp + c) * 6 + p];
  };

  if (want_tpt) {
    for (int b = 0; b < B; ++b) {
      if (!res[b].gradient || pb.st[b].status != RISM3D_OK) continue;
      hipLaunchKernelGGL(adj_coef, dim3(cp), dim3(NT), 0, pb.s, g, sites[b], pb.st[b].t64.p,
                         pb.st[b].bv64.p, (const double*)nullptr, pd.L[0], pd.L[1], pd.L[2],
                         pb.h[0], pb.h[1], pb.h[2], pd.rho, 0, A.p, Bf.p, (double*)nullptr,
                         (double*)nullptr);
      std::vector<double> acc;
      field_reduce(b, 0, acc);
      for (int a = 0; a < sites[b].n; ++a)
        for (int d = 0; d < 3; ++d) res[b].gradient[a * 3 + d] = acc[a * 6 + d];
    }
  }
  if (!want_lam) return;
  // adjoint solve for every solute that wants it, batched
  std::vector<DevBuf<double>> tsave(B);
  for (int b = 0; b < B; ++b) {
    auto& s = pb.st[b];
    const bool want = (wants_adj(b) || wants_resp(b)) && s.status == RISM3D_OK;
    s.active = want;
    if (!want) continue;
    s.Dt = DevBuf<double>(mP);
    s.gmu = DevBuf<double>(mP);
    tsave[b] = DevBuf<double>(mP);
    HIPCHK(hipMemcpyAsync(tsave[b].p, s.t64.p, mP * 8, hipMemcpyDeviceToDevice, pb.s));
    hipLaunchKernelGGL(adj_coef, dim3(cp), dim3(NT), 0, pb.s, g, sites[b], s.t64.p, s.bv64.p,
                       (const double*)nullptr, pd.L[0], pd.L[1], pd.L[2], pb.h[0], pb.h[1],
                       pb.h[2], pd.rho, 1, (double*)nullptr, (double*)nullptr, s.Dt.p, s.gmu.p);
    HIPCHK(hipMemsetAsync(s.t64.p, 0, mP * 8, pb.s));
    s.iters = s.iters32 = 0;
    s.best = INFINITY;
    s.since_best = 0;
    s.status = RISM3D_NONCONVERGED;
  }
  // The linear solve l = D K l + g: in FP64, or after a mixed-precision
  // solve by defect correction (FP32 passes on a correction around an FP64
  // base, FP64 residual at every rebase) to the same kind of tolerance,
  // relative to max|g|.
  Config cf = cf0;
  cf.mode = 1;
  cf.precision = cf0.precision == RISM3D_MIXED ? RISM3D_MIXED : RISM3D_FP64;
  cf.polish = 0;
  cf.solver = RISM3D_SOLVER_MDIIS;
  cf.lam = 0.5;
  cf.rel_tol = cf.precision == RISM3D_MIXED ? 1e-10 : 1e-11;
  iterate_batch(pb, cf);
  for (int b = 0; b < B; ++b) {
    auto& s = pb.st[b];
    if (!tsave[b].p) continue;
    res[b].adjoint_iterations = s.iters;
    const int n = sites[b].n;
    if (s.status != RISM3D_OK) {
      if (wants_adj(b))
        for (int q = 0; q < n * 6; ++q) res[b].dmu_dsite[q] = NAN;
      if (wants_resp(b))
        for (int q = 0; q < n * 3; ++q) res[b].response[q] = NAN;
      HIPCHK(hipMemcpyAsync(s.t64.p, tsave[b].p, mP * 8, hipMemcpyDeviceToDevice, pb.s));
      s.status = RISM3D_OK;
      continue;
    }
    // K lambda (mode 2) into R, then Lambda = r2c(lambda) (mode 3) into W
    std::vector<SoluteDesc> sd(B);
    for (auto& d : sd) std::memset(&d, 0, sizeof d);
    SoluteDesc& d = sd[b];
    d.active = 1;
    d.t = s.t64.p;
    d.r = s.R.p;
    d.unew = s.U.p;
    d.bv = s.bv64.p;
    d.W = s.W.p;
    d.Fs = s.fs64.p;
    d.part = s.part.p;
    d.mode = 2;
    evaluate(pb, false, sd);
    if (wants_resp(b)) {
      hipLaunchKernelGGL(adj_coef, dim3(cp), dim3(NT), 0, pb.s, g, sites[b], tsave[b].p,
                         s.bv64.p, (const double*)s.R.p, pd.L[0], pd.L[1], pd.L[2], pb.h[0],
                         pb.h[1], pb.h[2], pd.rho, 4, A.p, Bf.p, (double*)nullptr,
                         (double*)nullptr);
      std::vector<double> acc;
      field_reduce(b, 0, acc);
      for (int a = 0; a < n; ++a)
        for (int q = 0; q < 3; ++q) res[b].response[a * 3 + q] = acc[a * 6 + q];
    }
    if (!wants_adj(b)) {
      HIPCHK(hipMemcpyAsync(s.t64.p, tsave[b].p, mP * 8, hipMemcpyDeviceToDevice, pb.s));
      HIPCHK(hipStreamSynchronize(pb.s));
      s.status = RISM3D_OK;
      continue;
    }
    d.mode = 3;
    {
      // forward passes only: closure_r2c (identity), fft_y, zstar (forward)
      HIPCHK(hipMemcpyAsync(pb.d_sd.p, sd.data(), B * sizeof(SoluteDesc), hipMemcpyHostToDevice, pb.s));
      using T = double;
      constexpr int LN = y_lines<T>();
      const dim3 gx(resid_grid_x(g), g.m, B);
      const dim3 gy((g.NX + LN - 1) / LN, g.m * g.N2, B);
      launch_closure_r2c<T>(g.N0, gx, pb.s, pb.d_sd.p, g, pb.tw.d[1].p, pb.tw.d[0].p, hist_of(sd));
      launch_fft_y<T, false>(g.N1, gy, pb.s, pb.d_sd.p, g, pb.tw.d[2].p);
      launch_zstar<T>(g.N2, g.NX, g.N1, B, pb.s, pb.d_sd.p, g, pb.tw.d[3].p, pb.chi3.p, pb.tau3.p);
    }
    hipLaunchKernelGGL(adj_coef, dim3(cp), dim3(NT), 0, pb.s, g, sites[b], tsave[b].p, s.bv64.p,
                       (const double*)s.R.p, pd.L[0], pd.L[1], pd.L[2], pb.h[0], pb.h[1],
                       pb.h[2], pd.rho, 2, A.p, Bf.p, (double*)nullptr, (double*)nullptr);
    std::vector<double> acc;
    field_reduce(b, 1, acc);
    DevBuf<double> tpart((size_t)n * ck * 4);
    hipLaunchKernelGGL(grad_tau, dim3(ck, n), dim3(NT), 0, pb.s, g, sites[b],
                       (const cpx<double>*)s.W.p, pb.tau3.p, pb.chi3.p, tpart.p);
    std::vector<double> hp(tpart.n);
    HIPCHK(hipMemcpyAsync(hp.data(), tpart.p, hp.size() * 8, hipMemcpyDeviceToHost, pb.s));
    HIPCHK(hipMemcpyAsync(s.t64.p, tsave[b].p, mP * 8, hipMemcpyDeviceToDevice, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
    HIPCHK(hipGetLastError());
    for (int a = 0; a < n; ++a) {
      double tp[4] = {0, 0, 0, 0};
      for (int c = 0; c < ck; ++c)
        for (int p = 0; p < 4; ++p) tp[p] += hp[((size_t)a * ck + c) * 4 + p];
      for (int p = 0; p < 6; ++p) res[b].dmu_dsite[a * 6 + p] = acc[a * 6 + p];
      for (int p = 0; p < 4; ++p) res[b].dmu_dsite[a * 6 + p] += tp[p];
    }
    s.status = RISM3D_OK;
  }
}

std::mutex g_mutex;  // one solve per process at a time (shared streams)

// The 3D solvent-solvent solve (hnc3d_solvent_solve, hnc3d.c:804-1073) on
// the residual kernels with the m(m+1)/2 site-site pairs as the "sites":
// per pair the closure of beta v_short and the x/y FFTs, the per-k OZ
// solve in zstar (special 1), MDIIS on the pairs' t.  Post: mu by the m x m
// functional with the FFT representation of the long-range Coulomb
// (special 3), g, and chi - 1 at the corner (mode 10 + forward passes).
int solvent_run(const rism3d_problem* pd, int m, const rism3d_site* solvent,
                const rism3d_options& opt, const double* t_init, double* g_out,
                double* chi_out, double* t_out, rism3d_result* res) {
  if (!pd || !solvent || !res || m < 1 || m > MAX_SITES) return RISM3D_BAD_ARG;
  for (int d = 0; d < 3; ++d) {
    if (!supported_n(pd->N[d])) return RISM3D_UNSUPPORTED;
    if (!(pd->L[d] > 0)) return RISM3D_BAD_ARG;
  }
  if (pd->L[0] != pd->L[1] || pd->L[0] != pd->L[2]) return RISM3D_UNSUPPORTED;  // omega_intra
  if (!(pd->beta > 0) || !(pd->rho > 0)) return RISM3D_BAD_ARG;
  if (pd->closure < 0 || pd->closure > RISM3D_CLOSURE_PSE7) return RISM3D_BAD_ARG;
  if (opt.precision != RISM3D_FP64) return RISM3D_UNSUPPORTED;
  if (opt.solver != RISM3D_SOLVER_MDIIS && opt.solver != RISM3D_SOLVER_PICARD)
    return RISM3D_UNSUPPORTED;
  if (opt.diis_depth < 1 || opt.diis_depth > MAX_DIIS) return RISM3D_BAD_ARG;
  const int np = m * (m + 1) / 2;
  if ((size_t)np * zstar_lines<double>(pd->N[2], np) * lstride_for(pd->N[2]) * 16 > 64 * 1024)
    return RISM3D_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt.device) return RISM3D_NO_DEVICE;
  std::lock_guard<std::mutex> lock(g_mutex);
  struct EndCall {
    ~EndCall() { cache_end_call(); }
  } ec;
  HIPCHK(hipSetDevice(opt.device));
  Problem pb;
  pb.B = 1;
  pb.D = opt.solver == RISM3D_SOLVER_MDIIS ? opt.diis_depth : 1;
  pb.mixed = 0;
  for (int d = 0; d < 3; ++d) { pb.N[d] = pd->N[d]; pb.L[d] = pd->L[d]; pb.h[d] = pd->L[d] / pd->N[d]; }
  GridArgs& g = pb.g;
  std::memset(&g, 0, sizeof g);
  for (int i = 0; i < MAX_SITES; ++i) {
    g.sig[i] = 1.0;
    g.rw[i] = 1.0;
  }
  g.N0 = pd->N[0]; g.N1 = pd->N[1]; g.N2 = pd->N[2];
  g.NX = pd->N[0] / 2 + 1;
  g.NH = (g.NX + 15) / 16 * 16;
  g.m = np;
  g.msol = m;
  g.P = (int64_t)g.N0 * g.N1 * g.N2;
  g.K = (int64_t)g.NH * g.N1 * g.N2;
  g.dkx = 2 * M_PI / pd->L[0]; g.dky = 2 * M_PI / pd->L[1]; g.dkz = 2 * M_PI / pd->L[2];
  const double h3 = pb.h[0] * pb.h[1] * pb.h[2];
  const double L3 = pd->L[0] * pd->L[1] * pd->L[2];
  g.h3 = h3;
  g.inv_L3 = 1.0 / L3;
  g.a_scale = h3 / L3;
  g.beta = pd->beta;
  g.rho = pd->rho;
  g.renorm = 0;
  g.zorder = std::min(3, std::max(0, env_int("RISM3D_ZSTAR_ORDER", 3)));
  g.closure = pd->closure;
  g.pse_order = pd->closure == RISM3D_CLOSURE_KH ? 1 : (pd->closure >= RISM3D_CLOSURE_PSE0 ? pd->closure - RISM3D_CLOSURE_PSE0 : 0);
  for (int i = 0; i < m; ++i) g.q[i] = solvent[i].charge;
  pb.mP = (int64_t)np * g.P;
  HIPCHK(hipStreamCreateWithFlags(&pb.s, hipStreamNonBlocking));
  struct StreamGuard {  // idle before its buffers return to the cache
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
  } sg{pb.s};
  make_twiddles(pb.tw, pb.N, pb.s);

  // omega_ij(k) in the chi3 slots of the pairs
  pb.chi3 = DevBuf<double>((size_t)np * g.K);
  {
    std::vector<double> rij(np, 0.0);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < i; ++j) {
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d) d2 += (solvent[i].x[d] - solvent[j].x[d]) * (solvent[i].x[d] - solvent[j].x[d]);
        rij[i * (i + 1) / 2 + j] = std::sqrt(d2);
      }
    DevBuf<double> d_rij(np);
    HIPCHK(hipMemcpy(d_rij.p, rij.data(), np * 8, hipMemcpyHostToDevice));
    const int ck = (int)std::min<int64_t>(4096, (g.K + NT - 1) / NT);
    hipLaunchKernelGGL(omega3, dim3(ck, np), dim3(NT), 0, pb.s, g, pd->L[0], d_rij.p, pb.chi3.p);
    HIPCHK(hipStreamSynchronize(pb.s));
  }
  pb.st.resize(1);
  pb.nparts = resid_grid_x(g) * np;
  auto& s0 = pb.st[0];
  s0.U = DevBuf<double>((size_t)pb.D * pb.mP);
  s0.R = DevBuf<double>((size_t)pb.D * pb.mP);
  s0.t64 = DevBuf<double>(pb.mP);
  s0.bv64 = DevBuf<double>(pb.mP);
  s0.W = DevBuf<cpx<double>>((size_t)np * g.K);
  s0.part = DevBuf<double>((size_t)pb.nparts * NPART);
  pb.d_sd = DevBuf<SoluteDesc>(1);
  pb.d_out = DevBuf<double>(NPART);
  if (t_init)
    HIPCHK(hipMemcpyAsync(s0.t64.p, t_init, pb.mP * 8, hipMemcpyHostToDevice, pb.s));
  else
    HIPCHK(hipMemsetAsync(s0.t64.p, 0, pb.mP * 8, pb.s));
  {
    std::vector<double> sse(3 * m);
    for (int i = 0; i < m; ++i) {
      sse[3 * i] = solvent[i].sigma;
      sse[3 * i + 1] = solvent[i].epsilon;
      sse[3 * i + 2] = solvent[i].charge;
    }
    DevBuf<double> d_sse(sse.size());
    HIPCHK(hipMemcpy(d_sse.p, sse.data(), sse.size() * 8, hipMemcpyHostToDevice));
    const int cp = (int)std::min<int64_t>(4096, (g.P + NT - 1) / NT);
    hipLaunchKernelGGL(pair_field, dim3(cp, np), dim3(NT), 0, pb.s, g, d_sse.p, pd->L[0],
                       pd->L[1], pd->L[2], pb.h[0], pb.h[1], pb.h[2], s0.bv64.p);
    HIPCHK(hipStreamSynchronize(pb.s));
    HIPCHK(hipGetLastError());
  }

  Config cf;
  cf.solver = opt.solver;
  cf.precision = RISM3D_FP64;
  cf.polish = 0;
  cf.depth = pb.D;
  cf.verbose = opt.verbose;
  cf.max_iter = pd->max_iter > 0 ? pd->max_iter : 1000;
  cf.lam = opt.solver == RISM3D_SOLVER_PICARD ? pd->lambda : opt.diis_step;
  cf.tol = pd->norm_tol;
  cf.mixed_tol = opt.mixed_tol;
  pb.g.special = 1;
  iterate_batch(pb, cf);
  res->status = s0.status;
  res->iterations = s0.iters;
  res->iterations_fp32 = 0;
  res->residual = s0.resid;

  // u_L(r) = IFFT(U(k) s(k)) / L^3 into R slot 0 (one copy per pair)
  const int64_t P = g.P;
  std::vector<SoluteDesc> sd(1);
  std::memset(sd.data(), 0, sizeof(SoluteDesc));
  sd[0].active = 1;
  sd[0].t = s0.t64.p;
  sd[0].bv = s0.bv64.p;
  sd[0].W = s0.W.p;
  sd[0].part = s0.part.p;
  sd[0].r = s0.R.p;
  sd[0].mode = 2;
  pb.g.special = 3;
  evaluate(pb, false, sd);
  pb.g.special = 0;
  // mu and g
  {
    const int cp = (int)std::min<int64_t>(2048, (P + NT - 1) / NT);
    DevBuf<double> part((size_t)cp * 4);
    DevBuf<double> gbuf(g_out ? pb.mP : 0);
    hipLaunchKernelGGL(mu_pairs, dim3(cp), dim3(NT), 0, pb.s, g, s0.t64.p, s0.bv64.p,
                       s0.R.p, g_out ? gbuf.p : (double*)nullptr, part.p);
    std::vector<double> hp((size_t)cp * 4);
    HIPCHK(hipMemcpyAsync(hp.data(), part.p, hp.size() * 8, hipMemcpyDeviceToHost, pb.s));
    if (g_out) HIPCHK(hipMemcpyAsync(g_out, gbuf.p, pb.mP * 8, hipMemcpyDeviceToHost, pb.s));
    if (t_out) HIPCHK(hipMemcpyAsync(t_out, s0.t64.p, pb.mP * 8, hipMemcpyDeviceToHost, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
    double acc[4] = {0, 0, 0, 0};
    for (int c = 0; c < cp; ++c)
      for (int q = 0; q < 4; ++q) acc[q] += hp[(size_t)c * 4 + q];
    const double f = pd->rho * h3 / pd->beta;
    res->mu = f * acc[0];
    res->mu_hnc = f * acc[1];
    res->mu_kh = f * acc[2];
    res->mu_gf = f * acc[3];
    res->pmv_factor = NAN;
  }
  // chi - 1 at the corner: spectrum of rho h, forward passes only
  if (chi_out) {
    sd[0].mode = 10;
    HIPCHK(hipMemcpyAsync(pb.d_sd.p, sd.data(), sizeof(SoluteDesc), hipMemcpyHostToDevice, pb.s));
    using T = double;
    constexpr int LN = y_lines<T>();
    const dim3 gx(resid_grid_x(g), g.m, 1);
    const dim3 gy((g.NX + LN - 1) / LN, g.m * g.N2, 1);
    launch_closure_r2c<T>(g.N0, gx, pb.s, pb.d_sd.p, g, pb.tw.d[1].p, pb.tw.d[0].p, hist_of(sd));
    launch_fft_y<T, false>(g.N1, gy, pb.s, pb.d_sd.p, g, pb.tw.d[2].p);
    sd[0].mode = 3;
    HIPCHK(hipMemcpyAsync(pb.d_sd.p, sd.data(), sizeof(SoluteDesc), hipMemcpyHostToDevice, pb.s));
    launch_zstar<T>(g.N2, g.NX, g.N1, 1, pb.s, pb.d_sd.p, g, pb.tw.d[3].p, pb.chi3.p,
                    (const double*)nullptr);
    const int64_t Kv = (int64_t)g.NX * g.N1 * g.N2;
    DevBuf<double> cbuf((size_t)np * Kv * 2);
    const int ck = (int)std::min<int64_t>(4096, (g.K + NT - 1) / NT);
    hipLaunchKernelGGL(chi_from_h, dim3(ck, np), dim3(NT), 0, pb.s, g,
                       (const cpx<double>*)s0.W.p, pb.chi3.p, cbuf.p);
    HIPCHK(hipMemcpyAsync(chi_out, cbuf.p, cbuf.n * 8, hipMemcpyDeviceToHost, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
  }
  HIPCHK(hipGetLastError());
  return res->status;
}

// spec_out (a nested coarse solve): the forward spectra of the converged
// t, per solute (empty where it did not converge), moved out of the
// problem's buffers.
int run_impl(const Inputs& in, rism3d_result* res, std::vector<double>* t_init,
             std::vector<double>* t_final, std::vector<DevBuf<cpx<double>>>* spec_out) {
  int v = validate(in);
  if (v != RISM3D_OK) {
    for (int b = 0; b < in.B; ++b) res[b].status = v;
    return v;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= in.opt.device) {
    for (int b = 0; b < in.B; ++b) res[b].status = RISM3D_NO_DEVICE;
    return RISM3D_NO_DEVICE;
  }
  HIPCHK(hipSetDevice(in.opt.device));
  const rism3d_problem& pd = *in.pd;
  Problem pb;
  pb.B = in.B;
  // one residual or Jacobian application uses history slot 0 only
  pb.D = (in.opt.solver == RISM3D_SOLVER_MDIIS && !in.t_eval) ? in.opt.diis_depth : 1;
  pb.mixed = in.opt.precision != RISM3D_FP64;
  pb.noalloc = in.noalloc;
  if (in.noalloc) pb.use_graphs = 0;
  // BF16 residual history with defect correction only: the FP32-only study
  // iterates t itself to its floor, where the Gram rounding would show
  if (in.opt.precision != RISM3D_MIXED) pb.r16 = 0;
  if (!pb.r16) pb.u16 = 0;
  if (!pb.r16) pb.r8 = 0;
  for (int d = 0; d < 3; ++d) { pb.N[d] = pd.N[d]; pb.L[d] = pd.L[d]; pb.h[d] = pd.L[d] / pd.N[d]; }
  GridArgs& g = pb.g;
  std::memset(&g, 0, sizeof g);
  for (int i = 0; i < MAX_SITES; ++i) {
    g.sig[i] = 1.0;
    g.rw[i] = 1.0;
  }
  if (in.opt.site_rho)
    for (int i = 0; i < in.m; ++i) {
      g.rw[i] = in.opt.site_rho[i] / in.pd->rho;
      g.sig[i] = std::sqrt(g.rw[i]);
    }
  g.N0 = pd.N[0]; g.N1 = pd.N[1]; g.N2 = pd.N[2];
  g.NX = pd.N[0] / 2 + 1;
  g.NH = (g.NX + 15) / 16 * 16;
  g.m = in.m;
  g.nrad = pd.nrad;
  g.P = (int64_t)g.N0 * g.N1 * g.N2;
  g.K = (int64_t)g.NH * g.N1 * g.N2;
  g.dkx = 2 * M_PI / pd.L[0]; g.dky = 2 * M_PI / pd.L[1]; g.dkz = 2 * M_PI / pd.L[2];
  g.dk_tab = M_PI / pd.rmax;
  const double h3 = pb.h[0] * pb.h[1] * pb.h[2];
  const double L3 = pd.L[0] * pd.L[1] * pd.L[2];
  g.h3 = h3;
  g.inv_L3 = 1.0 / L3;
  g.a_scale = h3 / L3;
  g.tau_scale = pd.beta * EPSILON0INV / L3;
  g.beta = pd.beta;
  g.renorm = (in.opt.renorm && !in.opt.chi3d) ? 1 : 0;
  g.zorder = std::min(3, std::max(0, env_int("RISM3D_ZSTAR_ORDER", 3)));
  g.closure = pd.closure;
  g.pse_order = pd.closure == RISM3D_CLOSURE_KH ? 1 : (pd.closure >= RISM3D_CLOSURE_PSE0 ? pd.closure - RISM3D_CLOSURE_PSE0 : 0);
  for (int i = 0; i < in.m; ++i) g.q[i] = in.solvent[i].charge;
  pb.mP = (int64_t)in.m * g.P;

  if (in.stream)
    pb.s = in.stream;
  else
    HIPCHK(hipStreamCreateWithFlags(&pb.s, hipStreamNonBlocking));
  struct StreamGuard {  // idle before its buffers return to the cache
    hipStream_t s;
    bool own;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      if (own) (void)hipStreamDestroy(s);
    }
  } sg{pb.s, in.stream == nullptr};
  make_twiddles(pb.tw, pb.N, pb.s);

  // Solvent kernel: chi - delta and the renormalisation tables
  // (solvent_kernel_rism, hnc3d.c:1499-1566; rism_solvent_renorm,
  // rism.f90:141-208).  chi_k is C layout [m][m][nrad].
  const int m = in.m, nrad = pd.nrad;
  if (in.opt.chi3d) {
    // --solvent-3d: chi - 1 of a 3D solvent solve on this grid
    // (solvent_kernel_file, hnc3d.c:1434-1437, 1572-1590), no tau
    const size_t np = (size_t)m * (m + 1) / 2;
    pb.chi3 = DevBuf<double>(np * g.K);
    pb.chi3f = DevBuf<float>(np * g.K);
    pb.tau3 = DevBuf<double>((size_t)m * g.K);
    pb.tau3f = DevBuf<float>((size_t)m * g.K);
    HIPCHK(hipMemcpy2DAsync(pb.chi3.p, g.NH * 8, in.opt.chi3d, g.NX * 8, g.NX * 8,
                            np * g.N1 * g.N2, hipMemcpyHostToDevice, pb.s));
    const int ck = (int)std::min<int64_t>(4096, (np * g.K + NT - 1) / NT);
    hipLaunchKernelGGL(to_f32, dim3(ck), dim3(NT), 0, pb.s, pb.chi3f.p, pb.chi3.p,
                       (int64_t)(np * g.K));
    HIPCHK(hipMemsetAsync(pb.tau3.p, 0, (size_t)m * g.K * 8, pb.s));
    HIPCHK(hipMemsetAsync(pb.tau3f.p, 0, (size_t)m * g.K * 4, pb.s));
  } else {
  std::vector<double> chi_tab((size_t)m * (m + 1) / 2 * nrad), tau_tab((size_t)m * nrad);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k < nrad; ++k)
        chi_tab[((size_t)i * (i + 1) / 2 + j) * nrad + k] =  // S_ij = chi_ij sig_j / sig_i
            in.chi_k[((size_t)i * m + j) * nrad + k] * (g.sig[j] / g.sig[i]) -
            (i == j ? 1.0 : 0.0);
  for (int i = 0; i < m; ++i)
    for (int k = 0; k < nrad; ++k) {
      const double kk = (2 * k + 1) * g.dk_tab / 2;
      const double v = 4 * M_PI * std::exp(-kk * kk / (4 * G_COULOMB * G_COULOMB)) / (kk * kk);
      double x = 0.0;
      for (int j = 0; j < m; ++j) x += in.chi_k[((size_t)j * m + i) * nrad + k] * in.solvent[j].charge;
      tau_tab[(size_t)i * nrad + k] = x * v;
    }
  pb.chi_tab = DevBuf<double>(chi_tab.size());
  pb.tau_tab = DevBuf<double>(tau_tab.size());
  HIPCHK(hipMemcpyAsync(pb.chi_tab.p, chi_tab.data(), chi_tab.size() * 8, hipMemcpyHostToDevice, pb.s));
  HIPCHK(hipMemcpyAsync(pb.tau_tab.p, tau_tab.data(), tau_tab.size() * 8, hipMemcpyHostToDevice, pb.s));
  {
    const size_t np = (size_t)m * (m + 1) / 2;
    pb.chi3 = DevBuf<double>(np * g.K);
    pb.chi3f = DevBuf<float>(np * g.K);
    pb.tau3 = DevBuf<double>((size_t)m * g.K);
    pb.tau3f = DevBuf<float>((size_t)m * g.K);
    const int ck = (int)std::min<int64_t>(4096, (g.K + NT - 1) / NT);
    hipLaunchKernelGGL(ktab3, dim3(ck), dim3(NT), 0, pb.s, g, pb.chi_tab.p, pb.tau_tab.p,
                       pb.chi3.p, pb.chi3f.p, pb.tau3.p, pb.tau3f.p);
  }
  }

  // per-solute buffers
  const int B = in.B;
  pb.st.resize(B);
  pb.nparts = resid_grid_x(g) * m;
  const int64_t mK = (int64_t)m * g.K;
  std::vector<SitesDesc> sites(B);
  std::vector<double*> bv64p(B), tp(B), gp(B);
  std::vector<float*> bv32p(B);
  std::vector<cpx<double>*> fs64p(B);
  std::vector<cpx<float>*> fs32p(B);
  std::vector<std::vector<double>> site_host(B);
  for (int b = 0; b < B; ++b) {
    auto& s = pb.st[b];
    s.U = DevBuf<double>((size_t)pb.D * pb.mP);
    s.R = DevBuf<double>((size_t)pb.D * pb.mP);
    s.t64 = DevBuf<double>(pb.mP);
    if (pb.mixed) s.t32 = DevBuf<float>(pb.mP);
    s.bv64 = DevBuf<double>(pb.mP);
    if (pb.mixed) {
      s.bv32 = DevBuf<float>(pb.mP);
      s.r64 = DevBuf<double>(pb.mP);
      s.fb32 = DevBuf<float>(pb.mP);
      s.t32b = DevBuf<float>(pb.mP);
    }
    s.W = DevBuf<cpx<double>>(mK);
    s.fs64 = DevBuf<cpx<double>>(g.K);
    if (pb.mixed) s.fs32 = DevBuf<cpx<float>>(g.K + (int64_t)g.N1 * g.N2);  // fs_index
    s.part = DevBuf<double>((size_t)pb.nparts * NPART);
    s.n_sites = in.n_sites[b];
    auto& sh = site_host[b];
    sh.resize((size_t)s.n_sites * 6);
    for (int a = 0; a < s.n_sites; ++a) {
      const rism3d_site& st = in.solutes[b][a];
      sh[6 * a + 0] = st.x[0]; sh[6 * a + 1] = st.x[1]; sh[6 * a + 2] = st.x[2];
      sh[6 * a + 3] = st.sigma; sh[6 * a + 4] = st.epsilon; sh[6 * a + 5] = st.charge;
    }
    s.sites = DevBuf<double>(sh.size());
    HIPCHK(hipMemcpyAsync(s.sites.p, sh.data(), sh.size() * 8, hipMemcpyHostToDevice, pb.s));
    sites[b].xyz_s_e_q = s.sites.p;
    sites[b].n = s.n_sites;
    bv64p[b] = s.bv64.p;
    bv32p[b] = pb.mixed ? s.bv32.p : nullptr;
    fs64p[b] = s.fs64.p;
    fs32p[b] = pb.mixed ? s.fs32.p : nullptr;
    // initial iterate: zero (hnc3d.c:1845) or the restart token
    if (t_init && b == 0)
      HIPCHK(hipMemcpyAsync(s.t64.p, t_init->data(), pb.mP * 8, hipMemcpyHostToDevice, pb.s));
    else
      HIPCHK(hipMemsetAsync(s.t64.p, 0, pb.mP * 8, pb.s));
  }
  pb.d_sd = DevBuf<SoluteDesc>(B);
  pb.d_sites = DevBuf<SitesDesc>(B);
  pb.d_out = DevBuf<double>((size_t)B * NPART);
  pb.d_cdst = DevBuf<void*>(B);
  pb.d_csrc = DevBuf<void*>(B);
  pb.d_cmask = DevBuf<int>(B);
  HIPCHK(hipMemcpyAsync(pb.d_sites.p, sites.data(), B * sizeof(SitesDesc), hipMemcpyHostToDevice, pb.s));

  // setup: beta v (field0) and F(k) s(k)
  {
    std::vector<double> sse(3 * m);
    for (int i = 0; i < m; ++i) {
      sse[3 * i] = in.solvent[i].sigma;
      sse[3 * i + 1] = in.solvent[i].epsilon;
      sse[3 * i + 2] = in.solvent[i].charge;
    }
    DevBuf<double> d_sse(sse.size());
    DevBuf<double*> d_bv64(B);
    DevBuf<float*> d_bv32(B);
    DevBuf<cpx<double>*> d_fs64(B);
    DevBuf<cpx<float>*> d_fs32(B);
    HIPCHK(hipMemcpyAsync(d_sse.p, sse.data(), sse.size() * 8, hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipMemcpyAsync(d_bv64.p, bv64p.data(), B * sizeof(double*), hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipMemcpyAsync(d_bv32.p, bv32p.data(), B * sizeof(float*), hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipMemcpyAsync(d_fs64.p, fs64p.data(), B * sizeof(void*), hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipMemcpyAsync(d_fs32.p, fs32p.data(), B * sizeof(void*), hipMemcpyHostToDevice, pb.s));
    const int cp = (int)std::min<int64_t>(4096, (g.P + NT - 1) / NT);
    const int ck = (int)std::min<int64_t>(4096, (g.K + NT - 1) / NT);
    hipLaunchKernelGGL(g.m <= 3 ? field_bv<3> : field_bv<MAX_SITES>, dim3(cp, B), dim3(NT), 0, pb.s, pb.d_sites.p, g, d_sse.p,
                       pd.L[0], pd.L[1], pd.L[2], pb.h[0], pb.h[1], pb.h[2], d_bv64.p,
                       pb.mixed ? d_bv32.p : (float* const*)nullptr);
    hipLaunchKernelGGL(form_factor, dim3(ck, B), dim3(NT), 0, pb.s, pb.d_sites.p, g, d_fs64.p,
                       pb.mixed ? d_fs32.p : (cpx<float>* const*)nullptr);
    HIPCHK(hipStreamSynchronize(pb.s));
    HIPCHK(hipGetLastError());
  }

  // coarse start: the same problem on N/2 first, its t interpolated to N
  // (one level: starting the N/2 solve from N/4 as well measured no faster)
  bool coarse = in.opt.coarse_tol > 0 && !in.opt.chi3d && g.special == 0 && !t_init &&
                !spec_out && !in.j_dt && !in.t_eval;
  for (int d = 0; d < 3; ++d)
    coarse = coarse && pd.N[d] % 2 == 0 && supported_n(pd.N[d] / 2);
  if (coarse) {
    rism3d_problem pdc = pd;
    for (int d = 0; d < 3; ++d) pdc.N[d] = pd.N[d] / 2;
    pdc.norm_tol = in.opt.coarse_tol;
    Inputs ic = in;
    ic.pd = &pdc;
    ic.opt.coarse_tol = 0;
    ic.opt.polish_iters = 0;
    ic.opt.mixed_tol = std::max(in.opt.mixed_tol, in.opt.coarse_tol);
    std::vector<rism3d_result> rc(B);
    std::memset(rc.data(), 0, B * sizeof(rism3d_result));
    std::vector<DevBuf<cpx<double>>> spec;
    {
      struct ProfPause {  // kernel timing covers the grid N only
        bool saved = g_prof.on;
        ProfPause() { g_prof.on = false; }
        ~ProfPause() { g_prof.on = saved; }
      } pause;
      run_impl(ic, rc.data(), nullptr, nullptr, &spec);
    }
    const int N0c = pdc.N[0], N1c = pdc.N[1], N2c = pdc.N[2];
    const int NHc = (N0c / 2 + 1 + 15) / 16 * 16;
    const int64_t Kc = (int64_t)NHc * N1c * N2c;
    const double Pc = (double)N0c * N1c * N2c;
    std::vector<SoluteDesc> sd(B);
    std::memset(sd.data(), 0, B * sizeof(SoluteDesc));
    const int ck = (int)std::min<int64_t>(4096, (g.K + NT - 1) / NT);
    bool any = false;
    for (int b = 0; b < B; ++b) {
      if (b >= (int)spec.size() || !spec[b].p) continue;
      auto& s = pb.st[b];
      hipLaunchKernelGGL(spec_upsample, dim3(ck, m), dim3(NT), 0, pb.s, spec[b].p, s.W.p, g,
                         N0c, N1c, N2c, NHc, Kc, 1.0 / Pc);
      // inverse pass (zstar mode 4, y, c2r with r = z) into t
      sd[b].t = s.t64.p;
      sd[b].r = s.t64.p;
      sd[b].bv = s.bv64.p;
      sd[b].W = s.W.p;
      sd[b].Fs = s.fs64.p;
      sd[b].part = s.part.p;
      sd[b].mode = 4;
      sd[b].active = 1;
      any = true;
    }
    if (any) {
      upload_sd(pb, sd, pb.d_sd.p);
      using T = double;
      constexpr int LN = y_lines<T>();
      const dim3 gx(resid_grid_x(g), m, B);
      const dim3 gy((g.NX + LN - 1) / LN, m * g.N2, B);
      launch_zstar<T>(g.N2, g.NX, g.N1, B, pb.s, pb.d_sd.p, g, pb.tw.d[3].p, pb.chi3.p, pb.tau3.p);
      launch_fft_y<T, true>(g.N1, gy, pb.s, pb.d_sd.p, g, pb.tw.d[2].p);
      launch_c2r_resid<T>(g.N0, gx, pb.s, pb.d_sd.p, g, pb.tw.d[1].p, pb.tw.d[0].p, 0);
      HIPCHK(hipStreamSynchronize(pb.s));
      HIPCHK(hipGetLastError());
    }
  }

  if (in.j_dt) {
    // one Jacobian application at t (jacobian_t1, hnc3d.c:1290-1362)
    auto& s0 = pb.st[0];
    const int64_t mP = pb.mP;
    const int ch = (int)std::min<int64_t>(2048, (mP + NT - 1) / NT);
    DevBuf<double> D(mP), v(mP), jv(mP);
    HIPCHK(hipMemcpyAsync(s0.t64.p, in.t_eval, mP * 8, hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipMemcpyAsync(v.p, in.j_dt, mP * 8, hipMemcpyHostToDevice, pb.s));
    hipLaunchKernelGGL(closure_dcdt, dim3(ch), dim3(NT), 0, pb.s, g, s0.t64.p, s0.bv64.p, D.p, mP);
    std::vector<SoluteDesc> sd(1);
    std::memset(sd.data(), 0, sizeof(SoluteDesc));
    sd[0].t = v.p;
    sd[0].r = jv.p;
    sd[0].bv = s0.bv64.p;
    sd[0].W = s0.W.p;
    sd[0].Fs = s0.fs64.p;
    sd[0].part = s0.part.p;
    sd[0].dt = D.p;
    sd[0].mode = in.j_flag ? 7 : 6;
    sd[0].active = 1;
    evaluate(pb, false, sd);
    HIPCHK(hipMemcpyAsync(in.r_out, jv.p, mP * 8, hipMemcpyDeviceToHost, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
    HIPCHK(hipGetLastError());
    res[0].status = RISM3D_OK;
    return RISM3D_OK;
  }

  if (in.t_eval) {
    // one residual evaluation F(t) = T(t) - t (iterate_t1, hnc3d.c:1184)
    auto& s0 = pb.st[0];
    HIPCHK(hipMemcpyAsync(s0.t64.p, in.t_eval, pb.mP * 8, hipMemcpyHostToDevice, pb.s));
    if (pb.mixed) convert_iterates(pb, std::vector<int>(1, 1), false);
    std::vector<SoluteDesc> sd(1);
    std::memset(sd.data(), 0, sizeof(SoluteDesc));
    sd[0].t = pb.mixed ? (void*)s0.t32.p : (void*)s0.t64.p;
    sd[0].r = s0.R.p;
    sd[0].unew = s0.U.p;
    sd[0].bv = pb.mixed ? (const void*)s0.bv32.p : (const void*)s0.bv64.p;
    sd[0].W = s0.W.p;
    sd[0].Fs = pb.mixed ? (const void*)s0.fs32.p : (const void*)s0.fs64.p;
    sd[0].part = s0.part.p;
    sd[0].active = 1;
    evaluate(pb, pb.mixed != 0, sd);
    if (pb.mixed) {
      // FP32 residual to FP64 through the (now free) iterate buffers
      HIPCHK(hipMemcpyAsync(s0.t32.p, s0.R.p, pb.mP * 4, hipMemcpyDeviceToDevice, pb.s));
      convert_iterates(pb, std::vector<int>(1, 1), true);
      HIPCHK(hipMemcpyAsync(in.r_out, s0.t64.p, pb.mP * 8, hipMemcpyDeviceToHost, pb.s));
    } else {
      HIPCHK(hipMemcpyAsync(in.r_out, s0.R.p, pb.mP * 8, hipMemcpyDeviceToHost, pb.s));
    }
    HIPCHK(hipStreamSynchronize(pb.s));
    HIPCHK(hipGetLastError());
    res[0].status = RISM3D_OK;
    return RISM3D_OK;
  }

  // autotuner: the launch configuration of each working precision, on
  // this batch (t, the history and the spectra are not read after it
  // before the iteration writes them)
  if (in.opt.tune && !in.noalloc) {
    if (pb.mixed) tune(pb, true, in.opt.verbose);
    tune(pb, false, in.opt.verbose);
  }

  Config cf;
  cf.solver = in.opt.solver;
  cf.precision = in.opt.precision;
  cf.polish = in.opt.polish_iters;
  cf.depth = pb.D;
  cf.verbose = in.opt.verbose;
  cf.max_iter = pd.max_iter > 0 ? pd.max_iter : 1000;
  cf.lam = (in.opt.solver == RISM3D_SOLVER_PICARD || in.opt.solver == RISM3D_SOLVER_JAGER)
               ? pd.lambda : in.opt.diis_step;
  cf.tol = pd.norm_tol;
  cf.mixed_tol = in.opt.mixed_tol;
  if (cf.solver >= RISM3D_SOLVER_NEWTON) {
    reference_solver(pb, cf);
  } else {
    iterate_batch(pb, cf);
    if (in.opt.continuation) continuation(pb, cf, pd.closure);
  }

  // observables in FP64 from the converged t (hnc3d.c:2157-2268)
  {
    std::vector<const double*> tt(B);
    std::vector<double*> gg(B, nullptr);
    std::vector<DevBuf<double>> gbuf(B);
    bool want_g = false;
    for (int b = 0; b < B; ++b) {
      tt[b] = pb.st[b].t64.p;
      if (res[b].g_device) {  // the caller's device buffer
        gg[b] = res[b].g_device;
        want_g = true;
      } else if (res[b].g) {
        gbuf[b] = DevBuf<double>(pb.mP);
        gg[b] = gbuf[b].p;
        want_g = true;
      }
    }
    DevBuf<const double*> d_tt(B);
    DevBuf<double*> d_gg(B);
    DevBuf<const double*> d_bv(B);
    std::vector<const double*> bvc(bv64p.begin(), bv64p.end());
    HIPCHK(hipMemcpyAsync(d_tt.p, tt.data(), B * sizeof(void*), hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipMemcpyAsync(d_gg.p, gg.data(), B * sizeof(void*), hipMemcpyHostToDevice, pb.s));
    HIPCHK(hipMemcpyAsync(d_bv.p, bvc.data(), B * sizeof(void*), hipMemcpyHostToDevice, pb.s));
    const int cp = (int)std::min<int64_t>(2048, (g.P + NT - 1) / NT);
    DevBuf<double> d_mu((size_t)B * cp * 5);
    hipLaunchKernelGGL(mu_density, dim3(cp, B), dim3(NT), 0, pb.s, pb.d_sites.p, g, d_tt.p, d_bv.p,
                       pd.L[0], pd.L[1], pd.L[2], pb.h[0], pb.h[1], pb.h[2],
                       want_g ? d_gg.p : (double* const*)nullptr, d_mu.p);
    std::vector<double> mu((size_t)B * cp * 5);
    HIPCHK(hipMemcpyAsync(mu.data(), d_mu.p, mu.size() * 8, hipMemcpyDeviceToHost, pb.s));
    HIPCHK(hipStreamSynchronize(pb.s));
    HIPCHK(hipGetLastError());
    for (int b = 0; b < B; ++b) {
      double acc[5] = {0, 0, 0, 0, 0}, comp[5] = {0, 0, 0, 0, 0};
      for (int c = 0; c < cp; ++c)
        for (int q = 0; q < 5; ++q) {  // Neumaier, fixed order
          const double v = mu[((size_t)b * cp + c) * 5 + q];
          const double s = acc[q] + v;
          comp[q] += (std::fabs(acc[q]) >= std::fabs(v)) ? (acc[q] - s) + v : (v - s) + acc[q];
          acc[q] = s;
        }
      for (int q = 0; q < 5; ++q) acc[q] += comp[q];
      const double f = pd.rho * h3 / pd.beta;
      auto& s = pb.st[b];
      res[b].mu = f * acc[0];
      res[b].mu_hnc = f * acc[1];
      res[b].mu_kh = f * acc[2];
      res[b].mu_gf = f * acc[3];
      res[b].pmv_factor = 1.0 - acc[4] * pd.rho * h3;
      res[b].residual = s.resid;
      res[b].iterations = s.iters;
      res[b].iterations_fp32 = s.iters32;
      res[b].status = s.status;
      if (res[b].g)
        HIPCHK(hipMemcpy(res[b].g, gg[b], pb.mP * 8, hipMemcpyDeviceToHost));
      if (res[b].t)
        HIPCHK(hipMemcpy(res[b].t, tt[b], pb.mP * 8, hipMemcpyDeviceToHost));
      if (t_final && b == 0) {
        t_final->resize(pb.mP);
        HIPCHK(hipMemcpy(t_final->data(), tt[b], pb.mP * 8, hipMemcpyDeviceToHost));
      }
    }
  }
  if (spec_out) {
    // forward spectrum of each converged t (closure pass with c = t, y,
    // zstar mode 3), handed to the caller with its buffer
    std::vector<SoluteDesc> sd(B);
    std::memset(sd.data(), 0, B * sizeof(SoluteDesc));
    for (int b = 0; b < B; ++b) {
      auto& s = pb.st[b];
      sd[b].t = s.t64.p;
      sd[b].bv = s.bv64.p;
      sd[b].W = s.W.p;
      sd[b].Fs = s.fs64.p;
      sd[b].part = s.part.p;
      sd[b].mode = 3;
      sd[b].active = s.status == RISM3D_OK ? 1 : 0;
    }
    upload_sd(pb, sd, pb.d_sd.p);
    using T = double;
    constexpr int LN = y_lines<T>();
    const dim3 gx(resid_grid_x(g), m, B);
    const dim3 gy((g.NX + LN - 1) / LN, m * g.N2, B);
    launch_closure_r2c<T>(g.N0, gx, pb.s, pb.d_sd.p, g, pb.tw.d[1].p, pb.tw.d[0].p, 0);
    launch_fft_y<T, false>(g.N1, gy, pb.s, pb.d_sd.p, g, pb.tw.d[2].p);
    launch_zstar<T>(g.N2, g.NX, g.N1, B, pb.s, pb.d_sd.p, g, pb.tw.d[3].p, pb.chi3.p, pb.tau3.p);
    HIPCHK(hipStreamSynchronize(pb.s));
    HIPCHK(hipGetLastError());
    spec_out->resize(B);
    for (int b = 0; b < B; ++b)
      if (sd[b].active) (*spec_out)[b] = std::move(pb.st[b].W);
    return RISM3D_OK;
  }
  for (int b = 0; b < B; ++b)
    if (res[b].info && pb.st[b].status == RISM3D_OK) info_observables(pb, in, b, sites[b], res[b].info);
  gradients(pb, in, res, cf, sites);

  int worst = RISM3D_OK;
  for (int b = 0; b < B; ++b) worst = std::max(worst, res[b].status);
  return worst;
}

int run(const Inputs& in, rism3d_result* res, std::vector<double>* t_init,
        std::vector<double>* t_final) {
  std::lock_guard<std::mutex> lock(g_mutex);
  struct EndCall {  // an allocation-free call leaves the cache alone
    bool on;
    ~EndCall() {
      if (on) cache_end_call();
    }
  } ec{!in.noalloc};
  return run_impl(in, res, t_init, t_final, nullptr);
}

// Device bytes one residual evaluation takes from the arena (the buffers
// run_impl creates on the t_eval path with one history slot), each rounded
// up to the arena's 256-byte alignment; the residual-workgroup partials
// are bounded by one workgroup per (y, z) row and site.
size_t noalloc_bytes(const rism3d_problem& pd, int m, int n, bool mixed) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t N0 = pd.N[0], N1 = pd.N[1], N2 = pd.N[2];
  const size_t NH = (N0 / 2 + 1 + 15) / 16 * 16;
  const size_t P = N0 * N1 * N2, K = NH * N1 * N2, mP = (size_t)m * P;
  const size_t np = (size_t)m * (m + 1) / 2, nrad = (size_t)std::max(pd.nrad, 0);
  size_t b = 0;
  for (size_t len : {N0, N0 / 2, N1, N2}) b += al(len * 16) + al(len * 8);  // twiddles
  b += al(np * nrad * 8) + al(m * nrad * 8);                       // 1D tables
  b += al(np * K * 8) + al(np * K * 4) + al(m * K * 8) + al(m * K * 4);  // 3D tables
  b += 4 * al(mP * 8) + al(m * K * 16) + al(K * 16);  // U, R, t, beta v; W; F s
  if (mixed) b += 4 * al(mP * 4) + al(mP * 8) + al((K + N1 * N2) * 8);
  b += al(N1 * N2 * m * NPART * 8) + al((size_t)std::max(n, 0) * 6 * 8);
  b += al(sizeof(SoluteDesc)) + al(sizeof(SitesDesc)) + al(NPART * 8) + 3 * al(8);
  b += al(3 * (size_t)m * 8) + 4 * al(8);  // setup descriptors
  return b;
}

// Per device: the address of the fixed scratch and the stream of the
// allocation-free entry point, set up on its first call on that device.
struct NoallocDevice {
  std::mutex mu;
  std::map<int, std::pair<void*, hipStream_t>> dev;
};
NoallocDevice& noalloc_device() {
  static NoallocDevice d;
  return d;
}

}  // namespace
}  // namespace rism

struct rism3d_restart {
  int N[3];
  int m;
  std::vector<double> t;
};

extern "C" {

void rism3d_options_default(rism3d_options* opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof *opt);
  opt->renorm = 1;
  opt->solver = RISM3D_SOLVER_MDIIS;
  opt->precision = RISM3D_FP64;
  opt->polish_iters = 3;
  opt->mixed_tol = 1e-5;
  opt->diis_depth = 8;
  opt->diis_step = 0.5;
  opt->verbose = 0;
  opt->device = 0;
  opt->dielectric = 0.0;
  opt->continuation = 1;
}

int rism3d_solvent_chi_internal(const rism3d_problem* pd, int m, const rism3d_site* solvent,
                                double dielectric, double* chi_out);

// locked: take the library lock and trim the cache at the end (a lone
// call); otherwise the caller holds the lock (rism3d_batch_solve_grids).
static int solve_impl(const rism3d_problem* pd, int m, const rism3d_site* solvent, int B,
                      const int* n_sites, const rism3d_site* const* solutes, const double* chi_k,
                      const rism3d_options* opt, rism3d_result* res, rism3d_restart** restart,
                      bool locked = true) {
  if (!pd || !res || !n_sites || !solutes) return RISM3D_BAD_ARG;
  rism::Inputs in;
  in.pd = pd;
  in.m = m;
  in.solvent = solvent;
  in.B = B;
  in.n_sites = n_sites;
  in.solutes = solutes;
  if (opt) in.opt = *opt; else rism3d_options_default(&in.opt);
  std::vector<double> chi_own;
  rism3d_problem pd_up;
  if (in.opt.site_rho) {
    // a mixture brings its own chi (rism3d_solvent_chi_mixture); the
    // gradients are built for one density
    bool bad = !chi_k || in.opt.chi3d;
    for (int i = 0; i < m && !bad; ++i) bad = !(in.opt.site_rho[i] > 0);
    int st = bad ? RISM3D_BAD_ARG : RISM3D_OK;
    for (int b = 0; b < B && st == RISM3D_OK; ++b)
      if (res[b].gradient || res[b].dmu_dsite || res[b].response) st = RISM3D_UNSUPPORTED;
    if (st != RISM3D_OK) {
      for (int b = 0; b < B; ++b) res[b].status = st;
      return st;
    }
  }
  if (!chi_k && !in.opt.chi3d) {
    // chi from 1D DRISM on the upscaled radial grid (hnc3d.c:772-783,
    // 1471-1495): rmax = 4 max(L)/2, nrad = 16 max(N).
    pd_up = *pd;
    pd_up.rmax = 4 * std::max(std::max(pd->L[0], pd->L[1]), pd->L[2]) / 2;
    pd_up.nrad = 16 * std::max(std::max(pd->N[0], pd->N[1]), pd->N[2]);
    chi_own.resize((size_t)m * m * pd_up.nrad);
    int st = rism3d_solvent_chi_internal(&pd_up, m, solvent, in.opt.dielectric, chi_own.data());
    if (st != RISM3D_OK) {
      for (int b = 0; b < B; ++b) res[b].status = st;
      return st;
    }
    in.pd = &pd_up;
    chi_k = chi_own.data();
  }
  in.chi_k = chi_k;
  std::vector<double> t_init, t_final;
  std::vector<double>* ti = nullptr;
  if (restart && *restart) {
    rism3d_restart* r = *restart;
    if (r->m == m && r->N[0] == pd->N[0] && r->N[1] == pd->N[1] && r->N[2] == pd->N[2]) {
      t_init = std::move(r->t);
      ti = &t_init;
    }
    delete r;
    *restart = nullptr;
  }
  try {
    int st = locked ? rism::run(in, res, ti, restart ? &t_final : nullptr)
                    : rism::run_impl(in, res, ti, restart ? &t_final : nullptr, nullptr);
    if (restart && !t_final.empty()) {
      rism3d_restart* r = new rism3d_restart;
      for (int d = 0; d < 3; ++d) r->N[d] = pd->N[d];
      r->m = m;
      r->t = std::move(t_final);
      *restart = r;
    }
    return st;
  } catch (const rism::HipFailure& e) {
    const int st = (e.err == hipErrorOutOfMemory) ? RISM3D_OOM : RISM3D_NO_DEVICE;
    for (int b = 0; b < B; ++b) res[b].status = st;
    return st;
  }
}

int rism3d_solute_solve(const rism3d_problem* pd, int m, const rism3d_site* solvent, int n,
                        const rism3d_site* solute, const double* chi_k, const rism3d_options* opt,
                        rism3d_result* res, rism3d_restart** restart) {
  const rism3d_site* sol[1] = {solute};
  int ns[1] = {n};
  return solve_impl(pd, m, solvent, 1, ns, sol, chi_k, opt, res, restart);
}

int rism3d_batch_solve(const rism3d_problem* pd, int m, const rism3d_site* solvent, int n_solutes,
                       const int* n_sites, const rism3d_site* const* solutes, const double* chi_k,
                       const rism3d_options* opt, rism3d_result* res) {
  return solve_impl(pd, m, solvent, n_solutes, n_sites, solutes, chi_k, opt, res, nullptr);
}

int rism3d_batch_solve_grids(const rism3d_problem* pds, int m, const rism3d_site* solvent,
                             int n_solutes, const int* n_sites,
                             const rism3d_site* const* solutes, const double* chi_k,
                             const rism3d_options* opt, rism3d_result* res) {
  if (!pds || !res || !n_sites || !solutes || n_solutes < 1) return RISM3D_BAD_ARG;
  // one chi for all: the same radial grid
  for (int b = 1; b < n_solutes; ++b)
    if (chi_k && (pds[b].nrad != pds[0].nrad || pds[b].rmax != pds[0].rmax)) {
      for (int q = 0; q < n_solutes; ++q) res[q].status = RISM3D_BAD_ARG;
      return RISM3D_BAD_ARG;
    }
  // size classes: solutes whose problems are identical share one batch
  std::vector<std::vector<int>> groups;
  for (int b = 0; b < n_solutes; ++b) {
    bool placed = false;
    for (auto& gr : groups)
      if (std::memcmp(&pds[gr[0]], &pds[b], sizeof(rism3d_problem)) == 0) {
        gr.push_back(b);
        placed = true;
        break;
      }
    if (!placed) groups.push_back({b});
  }
  const int G = (int)groups.size();
  std::vector<std::vector<int>> ns(G);
  std::vector<std::vector<const rism3d_site*>> sol(G);
  std::vector<std::vector<rism3d_result>> rg(G);
  std::vector<int> st(G, RISM3D_OK);
  for (int k = 0; k < G; ++k)
    for (int b : groups[k]) {
      ns[k].push_back(n_sites[b]);
      sol[k].push_back(solutes[b]);
      rg[k].push_back(res[b]);
    }
  auto solve = [&](int k) {
    st[k] = solve_impl(&pds[groups[k][0]], m, solvent, (int)groups[k].size(), ns[k].data(),
                       sol[k].data(), chi_k, opt, rg[k].data(), nullptr, false);
  };
  {
    std::lock_guard<std::mutex> lock(rism::g_mutex);
    if (rism::g_prof.on || G == 1) {  // kernel timing records one stream at a time
      for (int k = 0; k < G; ++k) solve(k);
    } else {
      // every size class on its own stream, from its own host thread: the
      // classes' kernels share the device instead of running in turn
      std::vector<std::thread> th;
      for (int k = 0; k < G; ++k) th.emplace_back(solve, k);
      for (auto& t : th) t.join();
    }
    rism::cache_end_call();
  }
  int worst = RISM3D_OK;
  for (int k = 0; k < G; ++k) {
    for (size_t i = 0; i < groups[k].size(); ++i) res[groups[k][i]] = rg[k][i];
    if (st[k] != RISM3D_OK && (worst == RISM3D_OK || worst == RISM3D_NONCONVERGED)) worst = st[k];
  }
  return worst;
}

void rism3d_restart_destroy(rism3d_restart* restart) { delete restart; }

int rism3d_residual(const rism3d_problem* pd, int m, const rism3d_site* solvent, int n,
                    const rism3d_site* solute, const double* chi_k, const rism3d_options* opt,
                    const double* t, double* r) {
  if (!pd || !chi_k || !t || !r || !solute) return RISM3D_BAD_ARG;
  rism::Inputs in;
  in.pd = pd;
  in.m = m;
  in.solvent = solvent;
  in.B = 1;
  const rism3d_site* sol[1] = {solute};
  int ns[1] = {n};
  in.n_sites = ns;
  in.solutes = sol;
  in.chi_k = chi_k;
  if (opt) in.opt = *opt; else rism3d_options_default(&in.opt);
  in.t_eval = t;
  in.r_out = r;
  rism3d_result res;
  std::memset(&res, 0, sizeof res);
  try {
    return rism::run(in, &res, nullptr, nullptr);
  } catch (const rism::HipFailure& e) {
    return (e.err == hipErrorOutOfMemory) ? RISM3D_OOM : RISM3D_NO_DEVICE;
  }
}

int rism3d_residual_noalloc(const rism3d_problem* pd, int m, const rism3d_site* solvent, int n,
                            const rism3d_site* solute, const double* chi_k,
                            const rism3d_options* opt, const double* t, double* r) {
  if (!pd || !chi_k || !t || !r || !solute || !solvent || m < 1 || n < 1) return RISM3D_BAD_ARG;
  rism::Inputs in;
  in.pd = pd;
  in.m = m;
  in.solvent = solvent;
  in.B = 1;
  const rism3d_site* sol[1] = {solute};
  int ns[1] = {n};
  in.n_sites = ns;
  in.solutes = sol;
  in.chi_k = chi_k;
  if (opt) in.opt = *opt; else rism3d_options_default(&in.opt);
  in.opt.coarse_tol = 0;
  in.t_eval = t;
  in.r_out = r;
  const size_t cap = sizeof(scratch);
  const size_t need = rism::noalloc_bytes(*pd, m, n, in.opt.precision != RISM3D_FP64);
  if (need > cap) {
    std::fprintf(stderr,
                 "rism3d_residual_noalloc: this problem needs %zu bytes of device scratch, "
                 "the fixed scratch holds %zu (RISM3D_NOALLOC_SCRATCH_MB)\n", need, cap);
    return RISM3D_NO_SCRATCH;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || in.opt.device < 0 || ndev <= in.opt.device)
    return RISM3D_NO_DEVICE;
  rism3d_result res;
  std::memset(&res, 0, sizeof res);
  struct ArenaGuard {
    ~ArenaGuard() { rism::t_arena = nullptr; }
  } ag;
  try {
    void* base = nullptr;
    {
      rism::NoallocDevice& nd = rism::noalloc_device();
      std::lock_guard<std::mutex> lk(nd.mu);
      auto it = nd.dev.find(in.opt.device);
      if (it == nd.dev.end()) {
        if (hipSetDevice(in.opt.device) != hipSuccess) return RISM3D_NO_DEVICE;
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
          std::fprintf(stderr, "rism3d_residual_noalloc: hipGetSymbolAddress failed: %s\n",
                       hipGetErrorString(e));
          return RISM3D_NO_DEVICE;
        }
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
          return RISM3D_NO_DEVICE;
        it = nd.dev.emplace(in.opt.device, std::make_pair(p, st)).first;
      }
      base = it->second.first;
      in.stream = it->second.second;
    }
    rism::Arena arena;
    arena.base = static_cast<char*>(base);
    arena.size = cap;
    rism::t_arena = &arena;
    in.noalloc = 1;
    return rism::run(in, &res, nullptr, nullptr);
  } catch (const rism::ScratchFull&) {
    std::fprintf(stderr, "rism3d_residual_noalloc: the fixed device scratch is full\n");
    return RISM3D_NO_SCRATCH;
  } catch (const rism::HipFailure& e) {
    return (e.err == hipErrorOutOfMemory) ? RISM3D_OOM : RISM3D_NO_DEVICE;
  }
}

int rism3d_jacobian(const rism3d_problem* pd, int m, const rism3d_site* solvent, int n,
                    const rism3d_site* solute, const double* chi_k, const rism3d_options* opt,
                    const double* t, const double* dt, int flag, double* jdt) {
  if (!pd || !chi_k || !t || !dt || !jdt || !solute) return RISM3D_BAD_ARG;
  if (pd->closure == RISM3D_CLOSURE_PY) return RISM3D_UNSUPPORTED;
  rism::Inputs in;
  in.pd = pd;
  in.m = m;
  in.solvent = solvent;
  in.B = 1;
  const rism3d_site* sol[1] = {solute};
  int ns[1] = {n};
  in.n_sites = ns;
  in.solutes = sol;
  in.chi_k = chi_k;
  if (opt) in.opt = *opt; else rism3d_options_default(&in.opt);
  in.opt.precision = RISM3D_FP64;
  in.t_eval = t;
  in.j_dt = dt;
  in.j_flag = flag ? 1 : 0;
  in.r_out = jdt;
  rism3d_result res;
  std::memset(&res, 0, sizeof res);
  try {
    return rism::run(in, &res, nullptr, nullptr);
  } catch (const rism::HipFailure& e) {
    return (e.err == hipErrorOutOfMemory) ? RISM3D_OOM : RISM3D_NO_DEVICE;
  }
}

int rism3d_solvent_solve(const rism3d_problem* pd, int m, const rism3d_site* solvent,
                         const rism3d_options* opt, const double* t_init, double* g_out,
                         double* chi_out, double* t_out, rism3d_result* res) {
  rism3d_options o;
  if (opt) o = *opt; else rism3d_options_default(&o);
  try {
    const int st = rism::solvent_run(pd, m, solvent, o, t_init, g_out, chi_out, t_out, res);
    if (res && st != RISM3D_OK && st != RISM3D_NONCONVERGED && st != RISM3D_NAN) res->status = st;
    return st;
  } catch (const rism::HipFailure& e) {
    const int st = (e.err == hipErrorOutOfMemory) ? RISM3D_OOM : RISM3D_NO_DEVICE;
    if (res) res->status = st;
    return st;
  }
}

int rism3d_solvent_chi(const rism3d_problem* pd, int m, const rism3d_site* solvent,
                       double dielectric, double* chi_out) {
  return rism3d_solvent_chi_internal(pd, m, solvent, dielectric, chi_out);
}

int rism3d_cache_release(void) {
  std::lock_guard<std::mutex> lock(rism::g_mutex);
  rism::cache_empty(-1);
  return RISM3D_OK;
}

const char* rism3d_version(void) { return "rism3d-mi355x 0.1 (gfx950)"; }

int rism3d_profile_start(int device) {
  std::lock_guard<std::mutex> lock(rism::g_mutex);
  if (hipSetDevice(device) != hipSuccess) return RISM3D_NO_DEVICE;
  auto& pf = rism::g_prof;
  for (int k = 0; k < 5; ++k)
    for (int w = 0; w < 2; ++w)
      if (!pf.ev[k][w] && hipEventCreate(&pf.ev[k][w]) != hipSuccess) return RISM3D_NO_DEVICE;
  for (int k = 0; k < 2 * rism::PROF_KERNELS; ++k) {
    pf.ms[k] = 0;
    pf.bytes[k] = 0;
    pf.launches[k] = 0;
  }
  pf.on = true;
  return RISM3D_OK;
}

int rism3d_profile_read(double* ms, double* bytes, long* launches) {
  std::lock_guard<std::mutex> lock(rism::g_mutex);
  auto& pf = rism::g_prof;
  for (int k = 0; k < 2 * rism::PROF_KERNELS; ++k) {
    if (ms) ms[k] = pf.ms[k];
    if (bytes) bytes[k] = pf.bytes[k];
    if (launches) launches[k] = pf.launches[k];
  }
  pf.on = false;
  return RISM3D_OK;
}

}  // extern "C"
