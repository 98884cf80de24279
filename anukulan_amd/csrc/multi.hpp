// Solving a stack of independent LPs of uneven size, each with its own
// constraint matrix, in one set of launches. No counterpart in the reference,
// which solves LPs one at a time (solve_pdhg, src/anukulan/pdhg.cpp:377-1285).
//
// The LPs' scaled matrices are concatenated block-diagonally (K and K^T, one
// CSR each) and their vectors end to end; every period of fused steps is one
// launch pair over the whole stack, with each row block reading its own LP's
// step scalars. Each LP keeps its own host control (step size, primal
// weight, restarts, polishing, termination), so LP i takes the decisions its
// plain solve would; an LP that has stopped has its blocks return at once.
#pragma once

#include <vector>

#include "../../include/anukulan_hip.h"
#include "backend.hpp"
#include "lp.hpp"
#include "pdhg.hpp"

namespace anuk {

std::vector<SolveResult> solve_pdhg_multi(HipBackend& be, const std::vector<const StandardLp*>& lps,
                                          const anuk_pdhg_options& opt);

}  // namespace anuk
