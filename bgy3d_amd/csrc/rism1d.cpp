// 1D RISM / DRISM solvent susceptibility producer (host C++).
// Placeholder until the DRISM restatement lands: callers must pass chi_k.
#include "../../include/rism3d.h"

extern "C" int rism3d_solvent_chi_internal(const rism3d_problem* pd, int m,
                                           const rism3d_site* solvent, double dielectric,
                                           double* chi_out) {
  (void)pd; (void)m; (void)solvent; (void)dielectric; (void)chi_out;
  return RISM3D_UNSUPPORTED;
}
