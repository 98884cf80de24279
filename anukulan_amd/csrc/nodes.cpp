#include "nodes.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "batch.hpp"

namespace anuk {

namespace {

// Nodes handed to one batched solve: four groups of 64 lanes. The node LPs
// are materialised on the host (a copy of the working LP with the node's
// bounds) one chunk at a time, which bounds the host memory a large frontier
// takes.
constexpr Int kNodesPerChunk = 256;

}  // namespace

NodeRelaxation node_verdict(SolveResult solve) {
  NodeRelaxation out;
  out.proved_infeasible = solve.status == ANUK_PDHG_PRIMAL_INFEASIBLE;
  // A first-order point that is feasible but short of optimal has an
  // objective above the relaxation's optimum, so it bounds nothing; the dual
  // objective bounds the subtree by weak duality, provided the dual is
  // feasible to 1e-7 of its own size (branch_and_bound.cpp:536-564).
  const double dual_feasible_enough = 1e-7 * std::fmax(1.0, std::fabs(solve.residual.dual_objective));
  out.optimal = solve.status == ANUK_PDHG_OPTIMAL && solve.residual.dual_residual_inf <= dual_feasible_enough;
  if (out.optimal) out.bound = solve.residual.dual_objective;
  out.solve = std::move(solve);
  return out;
}

std::vector<NodeRelaxation> solve_node_relaxations(HipBackend& be, const StandardLp& lp,
                                                   const double* lower, const double* upper,
                                                   Int count, const anuk_pdhg_options& opt) {
  if (count < 0) throw std::invalid_argument("nodes: negative node count");
  if (count > 0 && (lower == nullptr || upper == nullptr))
    throw std::invalid_argument("nodes: null node bounds");
  if (opt.warm_x != nullptr || opt.warm_y != nullptr)
    throw std::invalid_argument("nodes: every node starts cold (a warm start would be the same for all)");
  const size_t n = uz(lp.cols());
  std::vector<NodeRelaxation> out;
  out.reserve(uz(count));
  const bool together = batch_supports(opt, nullptr);
  for (Int first = 0; first < count; first += kNodesPerChunk) {
    const Int chunk = std::min(kNodesPerChunk, count - first);
    std::vector<StandardLp> nodes(uz(chunk), lp);
    for (Int i = 0; i < chunk; ++i) {
      const size_t at = uz(first + i) * n;
      nodes[uz(i)].lower.assign(lower + at, lower + at + n);
      nodes[uz(i)].upper.assign(upper + at, upper + at + n);
    }
    if (together) {
      std::vector<const StandardLp*> ptrs;
      for (const StandardLp& s : nodes) ptrs.push_back(&s);
      for (SolveResult& r : solve_pdhg_batch(be, ptrs, opt)) out.push_back(node_verdict(std::move(r)));
    } else {
      for (const StandardLp& s : nodes) out.push_back(node_verdict(solve_pdhg(be, s, opt)));
    }
  }
  return out;
}

}  // namespace anuk
