// covfit.h -- the coverage-distribution stage (covfit.cpp) for the pipelined session.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/bsq.h"

namespace bsq {

// fit_coverage_distribution of one group's histogram n[0..N] (rows j >= 1) over the group's
// summed length; the 48 simplex restarts run on `threads` host threads (1: the calling thread).
void fit_coverage_group(const std::vector<double>& n, uint32_t N, uint32_t group_length, int threads,
                        bsq_coverage_fit& out, double& scale);
// analyze_unique_coverage_distributions over every coverage group of args.
void coverage_stage(const bsq_coverage_stage_args* args, bsq_coverage_summary* per_target, int threads);
// Mean of rows j >= 1 of h[0..len), summed as the fit sums it.
double coverage_histogram_mean(const double* h, int64_t len);

}  // namespace bsq
