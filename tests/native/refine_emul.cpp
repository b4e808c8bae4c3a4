// Host-side run of rustb_amd/csrc/refine_core.hpp (the cluster Rayleigh-Ritz
// executed per k-point by the GPU refinement kernel).  Test-only.
#include "../../rustb_amd/csrc/refine_core.hpp"

extern "C" void emul_cluster_rr(int n, double* Z, const double* HZ, double* E, double gap_rel) {
  static rtb::C2 buf[7][rtb::kMaxCluster][rtb::kMaxCluster];
  rtb::cluster_rr_one(n, reinterpret_cast<rtb::Cd*>(Z), reinterpret_cast<const rtb::Cd*>(HZ), E, gap_rel, buf);
}
