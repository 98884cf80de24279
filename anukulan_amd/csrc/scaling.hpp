// scale_lp (src/anukulan/scaling.cpp:32-100) on the device: the same Ruiz
// infinity-norm sweeps and the Pock-Chambolle one-norm sweep, each entry and
// vector updated with the same operations in the same order, so the scaled
// K, K^T, q, c, bounds and the diagonals D1, D2 carry exactly the bits the
// host's scale_lp produces. K and K^T cross PCIe once each way.
#pragma once

#include "backend.hpp"
#include "lp.hpp"

namespace anuk {

Scaling scale_lp_device(HipBackend& be, StandardLp* lp, int ruiz_iterations, bool pock_chambolle);

// Matrices at least this large are scaled on the device by the solvers.
inline constexpr Int kDeviceScalingNnz = Int{1} << 20;

}  // namespace anuk
