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

// Per-solute device pointers for one residual evaluation.  One entry per
// solute of a batch; blockIdx.z selects the entry.
struct SoluteDesc {
  const double* t;     // [m][N2][N1][N0] current iterate
  double* r;           // [m][N2][N1][N0] residual out, F(t) = T(t) - t
  const void* bv;      // [m][N2][N1][N0] beta*v_short, in the working type
  void* W;             // [m][N2][N1][NH] spectral work, cpx<working type>
  const void* Fs;      // [N2][N1][NH] form factor times s(k), cpx<working>
  double* part;        // partial (max|r|, sum r^2) per residual workgroup
  int active;          // 0: converged or failed, every workgroup returns
  int pad_;
};

// Grid and physics constants shared by a batch.
struct GridArgs {
  int N0, N1, N2, NH;  // NH = N0/2 + 1
  int m;               // solvent sites
  int nrad;            // 1D table length
  int64_t P, K;        // N0 N1 N2, NH N1 N2
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
  double q[MAX_SITES]; // solvent charges
};

// MDIIS dot products and updates, one entry per solute.
struct DiisDesc {
  const double* rnew;            // newest residual
  const double* rh[MAX_DIIS];    // residual history
  const double* th[MAX_DIIS];    // iterate history
  double coef[MAX_DIIS];         // DIIS coefficients
  double* tout;                  // next iterate
  double lam;                    // step applied to residuals
  int nh;                        // history length
  int active;
};

// Solute sites for the setup and mu kernels, per solute.
struct SitesDesc {
  const double* xyz_s_e_q;   // [n][6]: x, y, z, sigma, epsilon, charge
  int n;
  int pad_;
};

}  // namespace rism
