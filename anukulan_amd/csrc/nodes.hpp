// Branch-and-bound node relaxations, solved together.
//
// The reference's tree solves one node at a time; with the first-order node
// solver each node is solve_pdhg on the working LP with the node's column
// bounds (src/anukulan/branch_and_bound.cpp:528-566). Every node of a tree
// shares the working LP's constraint matrix and differs only in its bounds,
// which is the shape of a case stack: here the nodes are solved as one
// batch (solve_pdhg_batch), each node getting the answer its own solve_pdhg
// gives, and each node's verdict is the reference's: a bound only from the
// dual objective, and only when the dual is feasible enough to make it one.
#pragma once

#include <vector>

#include "../../include/anukulan_hip.h"
#include "backend.hpp"
#include "lp.hpp"
#include "pdhg.hpp"

namespace anuk {

struct NodeRelaxation {
  // The relaxation's dual objective is a proof of a lower bound on the
  // node's subtree (branch_and_bound.cpp:562-564).
  bool optimal = false;
  // The solve ended with a Farkas certificate (branch_and_bound.cpp:533).
  bool proved_infeasible = false;
  // residual.dual_objective when `optimal`, else -infinity (NodeSolution's
  // default, branch_and_bound.cpp:52).
  double bound = -kInf;
  SolveResult solve;
};

// Node i has column bounds lower[i*cols, (i+1)*cols) and upper[...] in place
// of lp's own; everything else is lp's. Options the batched loop runs
// (batch_supports) solve all nodes together; other options solve each node
// with the plain device solver, one after another.
std::vector<NodeRelaxation> solve_node_relaxations(HipBackend& be, const StandardLp& lp,
                                                   const double* lower, const double* upper,
                                                   Int count, const anuk_pdhg_options& opt);

// The reference's verdict on a first-order node solve
// (branch_and_bound.cpp:533-566), applied to a finished solve.
NodeRelaxation node_verdict(SolveResult solve);

}  // namespace anuk
