// Kernel argument structs shared by the device kernels and the host solver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rism {

constexpr int MAX_SITES = 8;      // solvent sites m
constexpr int MAX_DIIS = 16;      // MDIIS history
constexpr int NT = 256;           // threads per workgroup, every kernel

constexpr double EPSILON0INV = 331.84164;   // bgy3d.h:145
constexpr double G_COULOMB = 1.2;           // bgy3d-solvents.h:9

// Per-solute device pointers for one MDIIS iteration (residual evaluation
// fused with the DIIS combination and the Gram-matrix dot products).  One
// entry per solute of a batch.  Vectors are [m][N2][N1][N0] in the working
// type T of the launch (float in the FP32 phase, double in FP64).
struct SoluteDesc {
  void* t;                   // iterate: written (ncomb > 0) or read (ncomb == 0)
  const void* u[MAX_DIIS];   // u_j = t_j + lam r_j, the DIIS history
  double coef[MAX_DIIS];     // t = sum_j coef_j u_j
  int ncomb;                 // entries in the combination (0: t given)
  int ndot;                  // history residuals to dot the new one with
  const void* rh[MAX_DIIS];  // their residuals r_j
  void* r;                   // new residual out, F(t) = T(t) - t
  void* unew;                // new u = t + lam r out
  double lam;
  const void* bv;            // beta*v_short
  void* W;                   // [m][N2][N1][NH] spectral work, cpx<T>
  const void* Fs;            // [N2][N1][NH] form factor times s(k), cpx<T>
  double* part;              // per residual workgroup: max|r|, sum r^2, dots
  int active;                // 0: converged / failed / waiting: return
  int mode;                  // 0 F(t); 1 adjoint D K l + g - l; 2 K l; 3 spectrum;
                             // 5 defect correction: F(t_b + delta) for the
                             //   correction delta (t) around a base t_b, as
                             //   F(t_b) + K (c(t_b + delta) - c(t_b)) - delta;
                             // 6 J v = K (D v) - v; 7 J' v = K ((D + 1) v)
                             //   (D = dc/dt in dt, jacobian_t1 and its flag form)
  const double* dt;          // modes 1, 6, 7: D = dc/dt (FP64)
  const double* gv;          // mode 1: g = dmu/dt
  const void* fb;            // mode 5: F(t_b); bv then holds x_b = -beta v + t_b
                             //   (PY: exp(-beta v) - 1)
  const void* shift;         // mode 5: t = sum_j coef_j u_j - shift_w shift: the
  double shift_w;            //   entries of the previous base moved to this one
  int r16;                   // FP32: r and the history rh in BF16 (Gram only)
  int u16;                   // FP32: u history in difference form: u[0] the newest
                             //   u (FP32), u[j > 0] BF16 differences of consecutive
                             //   entries, t = sum_j coef_j u[j]; the residual pass
                             //   writes the new u over u[0] (unew) and its
                             //   difference from the old one to dnew
  void* dnew;                // u16: BF16 difference out
  int r8;                    // FP32 (with r16): r and rh in FP8 e4m3 instead of
                             //   BF16, each entry scaled into the e4m3 range
  float rsc;                 // r8: the new r is stored as r * rsc
  float rhinv[MAX_DIIS];     // r8: 1 / the scale of rh[j]
};
static_assert(sizeof(SoluteDesc) % 8 == 0, "SoluteDesc is copied in 8-byte words");
constexpr int NPART = 2 + MAX_DIIS;   // doubles per residual workgroup

// MDIIS state of one solute in the FP32 defect-correction phase with the
// difference-form u history and the FP8 residual history, kept on the
// device between passes (mdiis_step): the history bookkeeping the host loop
// would otherwise do between two residual passes.  The host takes a
// solute's state back when it asks for a rebase or fails.  gram is indexed
// by history slot, [slot][slot] with row stride MAX_DIIS.
struct MdiisDev {
  // fixed while the solute stays on the device
  float* ucur;            // the newest u (FP32)
  uint16_t* U;            // BF16 difference slots, mP each
  uint8_t* R;             // FP8 residual slots, mP each
  int64_t mP;
  double rb_rinf;         // max|F(t_b)| of the current base
  double tol_half;        // half the FP64-checked tolerance of the phase
  int D;                  // history depth of the phase
  int only32;             // FP32-only precision: no rebase on the 1e3 drop
  int max_iter;
  int base;               // rebases so far in this phase
  // state
  int n;                  // entries in order (oldest first)
  int slot_new;           // slot receiving the pending r and u
  int ndot;               // entries the pending r is dotted with (dotted)
  int since_best, stall32, iters, iters32;
  int hold;               // the 1e3 drop only marks the solute ready (flag 4),
                          // the phase's tolerance parks it (flag 5)
  int order[MAX_DIIS];
  int dotted[MAX_DIIS];
  int sbase[MAX_DIIS];    // base of each slot's entry
  float rsc[MAX_DIIS];    // FP8 scale of each slot's residual
  double coef[MAX_DIIS];  // t = sum coef_j u_{order_j} for the next pass
  double resid, best, best32;
  double gram[MAX_DIIS * MAX_DIIS];
};
static_assert(sizeof(MdiisDev) % 8 == 0, "MdiisDev is copied in 8-byte words");

// What mdiis_step reports to the host per solute and pass (pinned memory).
struct MdiisStatus {
  int flag;               // 0 next pass prepared, 1 rebase wanted, 2 not
                          // finite, 4 ready for a rebase, next pass prepared,
                          // 5 at the phase's tolerance: parked until a rebase
  int iters, iters32;
  int hist;               // flag 0: history slots of the next pass, max(ncomb, ndot)
  double resid, r2;       // max|F|, ||F||^2 of the pass
};

// FP8 residual history: the scale that puts max|r| = rinf at 240 (the
// largest e4m3 value is 448; the next pass's residual may grow 1.8x before
// it is clamped).
__host__ __device__ inline float fp8_scale(double rinf) {
  const double r = __builtin_isfinite(rinf) && rinf > 1e-30 ? rinf : 1.0;
  const double v = 240.0 / r;
  return (float)(v < 1e30 ? v : 1e30);
}
constexpr int INFO_NSUM = 22;         // grid sums of the info observables

// Grid and physics constants shared by a batch.
struct GridArgs {
  int N0, N1, N2;
  int NX;              // N0/2 + 1: kx points of the half-complex spectrum
  int NH;              // row stride of spectral arrays: NX rounded up to 16,
                       // so that kx tiles start on 128-byte lines
  int m;               // solvent sites (pairs in the solvent-solvent mode)
  int nrad;            // 1D table length
  int64_t P, K;        // N0 N1 N2, NH N1 N2 (K counts the padding)
  double dkx, dky, dkz;  // 2 pi / L
  double dk_tab;       // pi / rmax, 1D table spacing
  double a_scale;      // h^3 / L^3: forward h^3 and inverse 1/L^3 folded
  double tau_scale;    // beta / eps0 / L^3 (renorm)
  double beta;         // (no-renorm: beta q_j u_c terms)
  double inv_L3;       // 1 / L^3
  double h3;           // h^3
  int renorm;          // 1: tau, 0: Ng long-range form
  int closure;         // enum rism3d_closure
  int pse_order;       // n of PSE-n, 1 for KH
  int zorder;          // zstar grid order (see zstar_reg)
  int zB, ztiles;      // zorder 2: solutes and kx tiles of the launch
  int nyq;             // FP32 z pass: the Nyquist column in a launch of its own
  int zwide;           // FP32 z pass at N2 <= 128: 32 kx columns, 512 threads
  // ylay (the FP32 xy-plane chain): W is y-major, point (site, ky, z, kx <
  // N0/2) at site K + ky N2 NH + z NH + kx, so that the z pass reads whole
  // regions of the spectrum, and the Nyquist column kx = N0/2 is compact
  // behind the sites, (site, z, ky) at m K + site N1 N2 + z N1 + ky (cpx<T>
  // units).  0: W is z-major like the tables, [site][z][ky][NH].
  int ylay;
  int pad_ylay_;
  double q[MAX_SITES]; // solvent charges
  // zstar variants (0 solute residual): 1 solvent-solvent OZ on site-site
  // pairs (msol sites, pair p = i(i+1)/2 + j), 2 Poisson potential of a
  // charge density, 3 the unit long-range Coulomb u_L(k) s(k) / L^3
  int special;
  int msol;            // solvent sites of the solvent-solvent mode
  double rho;          // solvent number density
  // mixtures (per-site densities rho_i): sig_i = sqrt(rho_i / rho) scales the
  // symmetric table S = D^-1/2 chi D^1/2 back to h_i = sum_j c_j chi_ji,
  // rw_i = rho_i / rho weights mu; both 1 for one density
  double sig[MAX_SITES];
  double rw[MAX_SITES];
};

// Solute sites for the setup and mu kernels, per solute.
struct SitesDesc {
  const double* xyz_s_e_q;   // [n][6]: x, y, z, sigma, epsilon, charge
  int n;
  int pad_;
};

}  // namespace rism
