// fmm_cpu.h -- TEST INFRASTRUCTURE: the CPU oracle's fast multipole evaluation (fmm_cpu.cpp).
#pragma once

#include <vector>

namespace orc_fmm {

// pu[t] = sum_s qu[s] / |x_t - y_s| and pv likewise (no corrections, no 1/(4 pi)), by an octree FMM of
// order p whose leaves hold at most `leaf` sources on average; sources sxyz[3 S], targets txyz[3 T].
// Returns the leaf level.
int potentials(const std::vector<double> &sxyz, const std::vector<double> &qu, const std::vector<double> &qv,
               const std::vector<double> &txyz, int p, int leaf, std::vector<double> &pu, std::vector<double> &pv);

}  // namespace orc_fmm
