// Host-side interior-point driver over device-resident iterates: the
// reference's main loop (sdpa_main.cpp:294-377) with Newton::Mehrotra,
// StepLength, DirectionParameter, Phase, Solutions and Residuals, whose
// matrix work runs in the dd kernels and whose scalar logic runs here.
#pragma once
#include <string>
#include <vector>

#include "problem.hpp"
#include "sparse_schur.hpp"

namespace sdp {

enum PhaseType { noINFO, pFEAS, dFEAS, pdFEAS, pdINF, pFEAS_dINF, pINF_dFEAS, pdOPT, pUNBD, dUNBD };

struct IterRecord {
    int it;
    dd mu, thetaP, thetaD, objP, objD, alphaP, alphaD, beta;
};

struct Timers {
    double schur = 0, chol = 0, solve = 0, lanczos = 0, inverse = 0, gemm_other = 0, total = 0;
};

struct SolveOutput {
    int iterations = 0;
    int phase = noINFO, phase_reported = noINFO;
    dd mu{0, 0}, objValPrimal{0, 0}, objValDual{0, 0}, normPrimalVec{0, 0}, normDualMat{0, 0}, relgap{0, 0};
    std::vector<IterRecord> trace;
    std::string message;
    Timers t;
    double schur_ddflops = 0;  // algorithmic dd-flops of Schur assembly + factorisation, summed
    int refine_sweeps = 0;     // Ozaki forms: dd refinement sweeps kept
};

struct Context;  // device, stream and scratch of a library handle

// Solves the uploaded problem from lambdaStar I, or from *init; host
// problem needed for dimensions.  sparse: the sparse-Schur path's device
// image (null or off: the dense Schur path).
int solve(Context &ctx, const HostProblem &hp, DeviceProblem &dp, const Params &prm, SolveOutput &out,
          std::vector<dd> *xMat = nullptr, std::vector<dd> *yVec = nullptr, std::vector<dd> *zMat = nullptr,
          const InitPoint *init = nullptr, const sps::Device *sparse = nullptr);

// Truncated-Lanczos minimum eigenvalue estimate of L M L^T (one block).
dd lanczos_min_eigen(Context &ctx, int n, const dd *d_L, const dd *d_M);

}  // namespace sdp
