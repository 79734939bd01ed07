// Implicit differentiation of the interior-point solution at its final
// iterate: gradients of p* = b.y*, X* and y* with respect to b, C and the
// stored entries of the A_i, for a batch of cotangents.
//
// Internal convention (as the solver): min C.X s.t. A_i.X = b_i, X psd; dual
// max b.y s.t. sum_i y_i A_i + Z = C, Z psd.  The central-path equations
// A(X) = b, A^T(y) + Z = C, XZ = mu I linearised at the final (X, y, Z) with
// the HKM form of the complementarity row (the Newton system of the solver,
// Newton::compute_bMat_dense_SDP at sdpa_newton.cpp:979-1068 and
// compute_DxMat at :1329-1343) give, for perturbations (db, dC, dA):
//   Rd = dC - sum_i y_i dA_i,  rp_i = db_i - dA_i . X,
//   B dy = rp + A(X Rd Z^{-1}),  dZ = Rd - A^T(dy),  dX = -sym(X dZ Z^{-1}),
// with B_ij = A_i . (X A_j Z^{-1}) the Schur matrix at the final iterate.
// The transpose of this map for cotangents (Xbar, ybar, pbar) is
//   W = X sym(Xbar) Z^{-1},  yhat = ybar + pbar b + A(W),  v = B^{-1} yhat,
//   G = sym(X A^T(v) Z^{-1}) - sym(W),
//   bbar = v + pbar y,  Cbar = G,  Abar_i = -y_i G - v_i X.
// The map is the derivative of the central-path point only where the
// iterate lies on the path (XZ = mu I, where the HKM, AHO and NT forms
// coincide); the solver's last iterate is off-centre by a factor that does
// not shrink with mu, which biases the linearisation by the same relative
// amount.  Setup therefore first takes Newton steps toward the central point
// at mu = X.Z/n (the HKM centring direction with beta = 1, full steps),
// then assembles B at the centred point and factors it once; every
// cotangent reuses the factor, each solve with two sweeps of dd iterative
// refinement against B.
#pragma once
#include <string>
#include <vector>

#include "context.hpp"
#include "problem.hpp"

namespace sdp {

struct Adjoint {
    int m = 0, nB = 0;
    idx_t nn = 0;
    idx_t na = 0;  // stored A values (gradient entries per cotangent)
    int ntask = 0;
    dd *d_X = nullptr, *d_Zinv = nullptr, *d_y = nullptr, *d_B = nullptr, *d_L = nullptr;
    dd *d_Z = nullptr, *d_Rd = nullptr, *d_T1 = nullptr, *d_T2 = nullptr, *d_m1 = nullptr, *d_m2 = nullptr;
    double centrality = 0;  // max |XZ - mu I| / mu after centring
    int *d_tcon = nullptr;    // constraint of each inner-product task
    idx_t *d_aoff = nullptr;  // first gradient entry of each task
    // per-cotangent workspace (grown to the batch size)
    int kcap = 0;
    dd *d_S = nullptr, *d_P = nullptr, *d_W = nullptr, *d_V = nullptr;
    dd *d_yhat = nullptr, *d_v = nullptr, *d_r = nullptr;
    bool ready = false;

    // Upload the final iterate, centre it (center_steps Newton steps), form
    // Z^{-1}, assemble B and factor it.  Returns 0, > 0 when Z, X or B is not
    // positive definite (the Cholesky's info), < 0 on allocation or device
    // errors (err says which).
    int setup(Context &ctx, const HostProblem &hp, DeviceProblem &dp, const std::vector<dd> &X,
              const std::vector<dd> &y, const std::vector<dd> &Z, int center_steps, std::string &err);
    // K cotangents (device arrays): Xbar K x nn (blocks concatenated, each
    // n_l x n_l column-major), ybar K x m, pbar K (host).  Outputs (device):
    // bbar K x m, Cbar K x nn (symmetric, the gradient with respect to the
    // symmetric matrix C), Abar K x na in inner-product task order: per
    // constraint, per block of it, the stored entries of a sparse block (the
    // derivative with respect to the stored value, which sets both (r, c)
    // and (c, r)) or the n_l x n_l entries of a dense block (the gradient with
    // respect to the symmetric matrix).  Any output may be null.
    int apply(Context &ctx, DeviceProblem &dp, int K, const dd *Xbar, const dd *ybar, const dd *pbar, dd *bbar,
              dd *Cbar, dd *Abar, std::string &err);
    void release();
};

}  // namespace sdp
