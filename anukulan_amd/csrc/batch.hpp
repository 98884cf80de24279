// Batched PDHG: a stack of LPs that share the constraint matrix K (a base
// plan and its what-if cases), solved together on one device.
//
// K and K^T are uploaded once; scaling, ||K|| and the row-block plans are
// computed once for the whole stack, because they depend on K alone. Every
// per-LP vector lives row-interleaved with one wave64 lane per LP, so each
// sparse product reads the matrix once for all the LPs. Each LP keeps its
// own Control (step size, primal weight, Halpern epoch, restart and polish
// state, verdict), so every LP takes exactly the decisions a plain solve of
// it takes. An LP that stops is retired at that check: its lane goes idle and
// its iterate is copied out. Polishing sub-solves of the LPs that reach a
// polish checkpoint together run as one sub-batch.
#pragma once

#include <vector>

#include "../../include/anukulan_hip.h"
#include "backend.hpp"
#include "lp.hpp"
#include "pdhg.hpp"

namespace anuk {

// Solves every LP of `lps` (all with the same K, num_equalities, rows and
// columns) with one set of options. Results are in input order and match
// the plain solve of each LP bit for bit.
std::vector<SolveResult> solve_pdhg_batch(HipBackend& be, const std::vector<const StandardLp*>& lps,
                                          const anuk_pdhg_options& options);

// Whether the batched loop runs these options (the fused-loop set: constant
// step, Halpern or no restarts, fixed-point restart test, device products);
// *why receives the first reason it does not.
bool batch_supports(const anuk_pdhg_options& options, const char** why);

}  // namespace anuk
