// Handle types shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/rsc.h"
#include "device/factor.h"
#include "device/pcg.h"
#include "host/rsc_host.h"

struct rsc_factor {
  rsc::Csc A;
  std::vector<double> xyz;
  bool has_xyz = false;
  rsc::Plan plan;
  hipStream_t stream = nullptr;
  std::unique_ptr<rsc::FactorEngine> eng;
  rsc::DevCsr dA;
  rsc::PcgWork pcgw;
  rsc_factor_info info{};
  // diagonal rank constant the plan's ranks were built with; it persists
  // across numeric factorizations, as SparseSolver::_diagonalRankConstant
  // does (SparseSolver.cpp:174-218)
  double diag_const = 0.0;
};
