// Environment updates contract_L / contract_R as block triple products.
//
//   contract_L (tensors/DmrgContractions.h:34-158, matcher LAWA
//   DmrgIndexGymnastics.h:32-77):
//     Lnew[{Rout,Rin,Rmid}][b] += coeff_buildL * W(a,b) * Abra^T . Lold[qL][a] . Aket
//   contract_R (DmrgContractions.h:179-311, matcher AWAR :79-124):
//     Rnew[{Lout,Lin,Lmid}][a] += coeff_buildR * W(a,b) * Aket . Rold[qR][b] . Abra^T
//
// The planner enumerates the contributing terms in the reference's loop order,
// creates output blocks on first touch (as Tripod::push_back does) and hands
// the triple set to the device engine: X = Abra^T (L) or Aket (R), Y = the old
// environment, Z = Aket (L) or Abra^T (R).
#pragma once
#include <vector>

#include "pivot.hpp"

namespace vmps {

// Dense Biped: A[s] of an MPS site tensor.
struct DBiped {
  std::vector<Qn> in, out;
  std::vector<DenseBlock> block;
  std::unordered_map<Key2, size_t, Key2Hash> dict;
  std::vector<double> data;
};

struct EnvProblem {
  Sym sym;
  int which = 0;  // 0 = contract_L, 1 = contract_R
  std::vector<Qn> qloc, qop;
  Tripod old;
  std::vector<DBiped> Abra, Aket;                        // [s]
  std::vector<std::vector<std::vector<WBiped>>> W;       // [s1][s2][k]
  // result (filled by plan_env)
  std::vector<std::array<Qn, 3>> new_index;              // sectors in creation order
  std::vector<std::vector<int64_t>> new_block;           // [sector][a] -> output block id or -1
  std::vector<int32_t> new_adim;
};

// Build the triple set of the update; fills P.new_* with the output layout.
void plan_env(EnvProblem& P, TripleSet& S);

}  // namespace vmps
