// Adjoint of the interior-point solution (the map and its transpose are
// stated in adjoint.hpp).  The n x n products run in the dd GEMM, A(.) and
// A^T(.) in the problem's inner-product and plus kernels, the Schur solves
// in the dd triangular solves with refinement; the kernels here form the
// symmetric parts, the right-hand sides and the per-entry A gradients.
#include "adjoint.hpp"

#include <algorithm>

namespace sdp {

namespace {

__device__ __forceinline__ int block_of(idx_t e, int nB, const idx_t *boff) {
    int l = 0;
    while (l + 1 < nB && e >= boff[l + 1]) l++;
    return l;
}

// out_k = sym(in_k) over the blocks of K concatenated spaces (out may be in).
__global__ void k_sym_batch(idx_t nn, int nB, const idx_t *boff, const int *bsize, const dd *in, dd *out) {
    const idx_t e = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nn) return;
    const int l = block_of(e, nB, boff);
    const idx_t o = boff[l], n = bsize[l], q = e - o, i = q % n, j = q / n;
    if (i > j) return;
    const idx_t kb = (idx_t)blockIdx.y * nn;
    const dd a = in[kb + o + i + n * j], b = in[kb + o + j + n * i];
    const dd v = dd_mul_d(dd_add(a, b), 0.5);
    out[kb + o + i + n * j] = v;
    out[kb + o + j + n * i] = v;
}

// G_k = sym(T_k) - sym(W_k)
__global__ void k_sym_diff_batch(idx_t nn, int nB, const idx_t *boff, const int *bsize, const dd *T, const dd *W,
                                 dd *G) {
    const idx_t e = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nn) return;
    const int l = block_of(e, nB, boff);
    const idx_t o = boff[l], n = bsize[l], q = e - o, i = q % n, j = q / n;
    if (i > j) return;
    const idx_t kb = (idx_t)blockIdx.y * nn, pij = kb + o + i + n * j, pji = kb + o + j + n * i;
    const dd t = dd_add(T[pij], T[pji]), w = dd_add(W[pij], W[pji]);
    const dd v = dd_mul_d(dd_sub(t, w), 0.5);
    G[pij] = v;
    G[pji] = v;
}

// y := y + a + s*b (s from the host, per cotangent k)
__global__ void k_add2(int m, dd *y, const dd *a, dd s, const dd *b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    dd r = y[i];
    if (a) r = dd_add(r, a[i]);
    if (!dd_is_zero(s)) r = dd_add(r, dd_mul(s, b[i]));
    y[i] = r;
}

// Abar entries of one cotangent: task t (constraint i, block l), entry q:
// -(y_i G_rc + v_i X_rc), doubled for an off-diagonal stored entry of a
// sparse block (the stored value sets (r, c) and (c, r)).
__global__ void k_abar(int ntask, const int *tcon, const int *type, const int *block, const int *nnz, const idx_t *off,
                       const idx_t *aoff, const int *row, const int *col, const idx_t *boff, const int *bsize,
                       const dd *y, const dd *v, const dd *G, const dd *X, dd *Abar) {
    const int t = blockIdx.x;
    const int i = tcon[t], l = block[t], nz = nnz[t];
    const idx_t n = bsize[l], o = boff[l], so = off[t], ao = aoff[t];
    const dd yi = y[i], vi = v[i];
    for (int q = threadIdx.x; q < nz; q += blockDim.x) {
        idx_t r, c;
        bool twice = false;
        if (type[t] == 0) {
            r = row[so + q];
            c = col[so + q];
            twice = r != c;
        } else {
            r = q % n;
            c = q / n;
        }
        const idx_t p = o + r + n * c;
        dd a = dd_neg(dd_add(dd_mul(yi, G[p]), dd_mul(vi, X[p])));
        if (twice) a = dd_mul_d(a, 2.0);
        Abar[ao + q] = a;
    }
}

template <typename T>
bool grow(T *&p, size_t count) {
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(count, 1)) == hipSuccess;
}

}  // namespace

void Adjoint::release() {
    for (void *p : {(void *)d_X, (void *)d_Zinv, (void *)d_y, (void *)d_B, (void *)d_L, (void *)d_Z, (void *)d_Rd,
                    (void *)d_T1, (void *)d_T2, (void *)d_m1, (void *)d_m2, (void *)d_tcon, (void *)d_aoff, (void *)d_S,
                    (void *)d_P, (void *)d_W, (void *)d_V, (void *)d_yhat, (void *)d_v, (void *)d_r})
        if (p) (void)hipFree(p);
    d_X = d_Zinv = d_y = d_B = d_L = d_Z = d_Rd = d_T1 = d_T2 = d_m1 = d_m2 = nullptr;
    d_tcon = nullptr;
    d_aoff = nullptr;
    d_S = d_P = d_W = d_V = d_yhat = d_v = d_r = nullptr;
    kcap = 0;
    ready = false;
}

namespace {
// y := alpha*y + beta*x (elementwise, dd)
__global__ void k_axpby(idx_t len, dd alpha, dd *y, dd beta, const dd *x) {
    const idx_t i = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) y[i] = dd_add(dd_mul(alpha, y[i]), dd_mul(beta, x[i]));
}
void axpby(hipStream_t s, idx_t len, dd alpha, dd *y, dd beta, const dd *x) {
    if (len > 0) k_axpby<<<(unsigned)((len + 255) / 256), 256, 0, s>>>(len, alpha, y, beta, x);
}
}  // namespace

int Adjoint::setup(Context &ctx, const HostProblem &hp, DeviceProblem &dp, const std::vector<dd> &X,
                   const std::vector<dd> &y, const std::vector<dd> &Z, int center_steps, std::string &err) {
    release();
    m = hp.m;
    nB = (int)hp.bs.size();
    nn = hp.nn;
    if (hp.nlp > 0) {
        err = "adjoint: problems with LP entries are not supported";
        return -1;
    }
    if ((idx_t)X.size() != nn || (idx_t)Z.size() != nn || (int)y.size() != m) {
        err = "adjoint: no solution to differentiate (solve first)";
        return -1;
    }
    hipStream_t s = ctx.stream;
    // tasks in upload order (constraint i, then its blocks) and the offsets
    // of their gradient entries
    std::vector<int> tcon;
    std::vector<idx_t> aoff;
    na = 0;
    for (int i = 0; i < m; i++)
        for (size_t k = 0; k < hp.A[i].index.size(); k++) {
            const Block &b = hp.A[i].blk[k];
            tcon.push_back(i);
            aoff.push_back(na);
            na += b.dense ? (idx_t)b.n * b.n : (idx_t)b.nnz_count;
        }
    ntask = (int)tcon.size();
    if (ntask != dp.ntask) {
        err = "adjoint: task table mismatch";
        return -1;
    }
    if (!grow(d_X, nn) || !grow(d_Z, nn) || !grow(d_Zinv, nn) || !grow(d_Rd, nn) || !grow(d_T1, nn) ||
        !grow(d_T2, nn) || !grow(d_S, nn) || !grow(d_P, nn) || !grow(d_y, m) || !grow(d_m1, m) || !grow(d_m2, m) ||
        !grow(d_r, m) || !grow(d_B, (size_t)m * m) || !grow(d_L, (size_t)m * m) || !grow(d_tcon, ntask) ||
        !grow(d_aoff, ntask)) {
        err = "adjoint: device allocation failed";
        release();
        return -2;
    }
    (void)hipMemcpyAsync(d_X, X.data(), sizeof(dd) * nn, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_Z, Z.data(), sizeof(dd) * nn, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_y, y.data(), sizeof(dd) * m, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_tcon, tcon.data(), sizeof(int) * ntask, hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(d_aoff, aoff.data(), sizeof(idx_t) * ntask, hipMemcpyHostToDevice, s);
    const dd one = mk(1.0), zero = mk(0.0), mone = mk(-1.0);
    auto info_of = [&]() -> int64_t {
        int64_t info = 0;
        (void)hipMemcpyAsync(&info, ctx.d_info, sizeof(info), hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) return -1;
        return info;
    };
    // Z^{-1} = L_Z^{-T} L_Z^{-1} per block (Jal::getInvChol's factorisation)
    auto inverse_z = [&]() -> int {
        for (int l = 0; l < nB; l++) {
            const idx_t n = hp.bs[l], o = hp.boff[l];
            sdk::copy(s, n * n, d_Z + o, d_S);
            sdk::potrf_lower(s, n, d_S, n, ctx.d_info, ctx.d_rinv);
            const int64_t info = info_of();
            if (info != 0) {
                err = info < 0 ? "adjoint: device error" : "adjoint: Z is not positive definite";
                return info < 0 ? -3 : (int)info;
            }
            sdk::zero_upper(s, n, d_S);
            sdk::inv_lower(s, n, d_S, n, d_P, n);
            sdk::trmm_lltn_self(s, n, d_P, n, d_Zinv + o, n);
        }
        return 0;
    };
    // B at (X, Z^{-1}) and its Cholesky factor
    auto factor_b = [&]() -> int {
        schur(s, dp, d_X, d_Zinv, d_B, ctx.dot_T);
        sdk::copy(s, (idx_t)m * m, d_B, d_L);
        sdk::potrf_lower(s, m, d_L, m, ctx.d_info, ctx.d_rinv);
        const int64_t info = info_of();
        if (info != 0) {
            err = info < 0 ? "adjoint: device error" : "adjoint: Schur matrix is not positive definite";
            return info < 0 ? -3 : (int)info;
        }
        return 0;
    };
    // x := B^{-1} x with two refinement sweeps (r: residual scratch)
    auto solve_b = [&](const dd *rhs, dd *x, dd *r) {
        sdk::copy(s, m, rhs, x);
        ctx.trsv(s, 'N', m, d_L, m, x);
        ctx.trsv(s, 'T', m, d_L, m, x);
        for (int sweep = 0; sweep < 2; sweep++) {
            sdk::copy(s, m, rhs, r);
            sdk::gemv_n(s, m, m, mone, d_B, m, x, one, r);
            ctx.trsv(s, 'N', m, d_L, m, r);
            ctx.trsv(s, 'T', m, d_L, m, r);
            sdk::axpy(s, m, one, r, x);
        }
    };
    // ret_l = a_l * b_l * c_l per block (tmp: d_P)
    auto triple = [&](dd *ret, const dd *a, const dd *b, const dd *c) {
        for (int l = 0; l < nB; l++) {
            const idx_t n = hp.bs[l], o = hp.boff[l];
            sdk::gemm(s, 'N', n, n, n, one, a + o, n, b + o, n, zero, d_P, n);
            sdk::gemm(s, 'N', n, n, n, one, d_P, n, c + o, n, zero, ret + o, n);
        }
    };
    // mu = X.Z / n
    idx_t ndim = 0;
    for (int nb : hp.bs) ndim += nb;
    ctx.dot(nn, d_X, d_Z, 0);
    dd muv;
    (void)hipMemcpyAsync(&muv, ctx.d_scal, sizeof(dd), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) {
        err = "adjoint: device error";
        return -3;
    }
    const dd mu = dd_div_d(muv, (double)ndim);
    const double center_tol = 1e-10;  // full Newton steps converge quadratically; stop once centred
    // max |XZ - mu I| / mu (T1, T2 as scratch); < 0 on a device error
    auto measure = [&]() -> double {
        for (int l = 0; l < nB; l++) {
            const idx_t n = hp.bs[l], o = hp.boff[l];
            sdk::gemm(s, 'N', n, n, n, one, d_X + o, n, d_Z + o, n, zero, d_T1 + o, n);
            sdk::set_identity(s, n, dd_neg(mu), d_T2 + o);
        }
        sdk::axpy(s, nn, one, d_T2, d_T1);
        sdk::max_abs(s, nn, d_T1, ctx.d_scal + 1, ctx.d_max_work, ctx.d_counter + 1);
        dd dev;
        (void)hipMemcpyAsync(&dev, ctx.d_scal + 1, sizeof(dd), hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) return -1.0;
        return dev.hi / mu.hi;
    };
    int rc = inverse_z();
    if (rc != 0) return rc;
    centrality = measure();
    for (int it = 0; it < center_steps && centrality > center_tol; it++) {
        // residuals: Rd = C - Z - A^T(y), rp = b - A(X)
        plus_A(s, dp, d_y, d_Rd);
        sdk::scal(s, nn, mone, d_Rd);
        plus_C(s, dp, d_Rd);
        sdk::axpy(s, nn, mone, d_Z, d_Rd);
        inner_A(s, dp, d_X, d_m1, ctx.dot_T);
        axpby(s, m, mone, d_m1, one, dp.d_b);
        // T1 = mu Z^{-1} - X, T2 = X Rd Z^{-1};  rhs = rp - A(T1) + A(T2)
        sdk::copy(s, nn, d_Zinv, d_T1);
        axpby(s, nn, mu, d_T1, mone, d_X);
        triple(d_T2, d_X, d_Rd, d_Zinv);
        inner_A(s, dp, d_T1, d_m2, ctx.dot_T);
        sdk::axpy(s, m, mone, d_m2, d_m1);
        inner_A(s, dp, d_T2, d_m2, ctx.dot_T);
        sdk::axpy(s, m, one, d_m2, d_m1);
        if ((rc = factor_b()) != 0) return rc;
        solve_b(d_m1, d_m2, d_r);  // dy
        // dZ = Rd - A^T(dy);  dX = sym(T1 - X dZ Z^{-1})
        plus_A(s, dp, d_m2, d_S);
        sdk::axpy(s, nn, mone, d_S, d_Rd);  // Rd := dZ
        triple(d_T2, d_X, d_Rd, d_Zinv);
        sdk::axpy(s, nn, mone, d_T2, d_T1);  // T1 := T1 - X dZ Z^{-1}
        const dim3 eg((unsigned)((nn + 255) / 256), 1);
        k_sym_batch<<<eg, 256, 0, s>>>(nn, nB, dp.d_boff, dp.d_bsize, d_T1, d_T1);
        sdk::axpy(s, nn, one, d_T1, d_X);
        sdk::axpy(s, m, one, d_m2, d_y);
        sdk::axpy(s, nn, one, d_Rd, d_Z);
        if ((rc = inverse_z()) != 0) return rc;
        centrality = measure();
    }
    if (centrality < 0) {
        err = "adjoint: device error";
        return -3;
    }
    if ((rc = factor_b()) != 0) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) {
        err = "adjoint: device error";
        return -3;
    }
    ready = true;
    return 0;
}

int Adjoint::apply(Context &ctx, DeviceProblem &dp, int K, const dd *Xbar, const dd *ybar, const dd *pbar, dd *bbar,
                   dd *Cbar, dd *Abar, std::string &err) {
    if (!ready) {
        err = "adjoint: not set up";
        return -1;
    }
    if (K < 1) return 0;
    hipStream_t s = ctx.stream;
    if (K > kcap) {
        kcap = 0;
        if (!grow(d_S, (size_t)K * nn) || !grow(d_P, (size_t)K * nn) || !grow(d_W, (size_t)K * nn) ||
            !grow(d_V, (size_t)K * nn) || !grow(d_yhat, (size_t)K * m) || !grow(d_v, (size_t)K * m) ||
            !grow(d_r, (size_t)K * m)) {
            err = "adjoint: device allocation failed";
            release();
            return -2;
        }
        kcap = K;
    }
    const dim3 eg((unsigned)((nn + 255) / 256), (unsigned)K);
    const dd one = mk(1.0), zero = mk(0.0);
    // W_k = X sym(Xbar_k) Z^{-1}
    if (Xbar) {
        k_sym_batch<<<eg, 256, 0, s>>>(nn, nB, dp.d_boff, dp.d_bsize, Xbar, d_S);
        for (int l = 0; l < nB; l++) {
            const idx_t n = dp.bs[l], o = dp.boff[l];
            sdk::gemm(s, 'N', n, n, n, one, d_X + o, n, d_S + o, n, zero, d_P + o, n, K, 0, nn, nn);
            sdk::gemm(s, 'N', n, n, n, one, d_P + o, n, d_Zinv + o, n, zero, d_W + o, n, K, nn, 0, nn);
        }
    } else {
        sdk::fill(s, (idx_t)K * nn, zero, d_W);
    }
    // yhat_k = A(W_k) + ybar_k + pbar_k b;  v_k = B^{-1} yhat_k
    for (int k = 0; k < K; k++) {
        dd *yh = d_yhat + (idx_t)k * m, *v = d_v + (idx_t)k * m, *r = d_r + (idx_t)k * m;
        inner_A(s, dp, d_W + (idx_t)k * nn, yh, ctx.dot_T);
        k_add2<<<(m + 255) / 256, 256, 0, s>>>(m, yh, ybar ? ybar + (idx_t)k * m : nullptr, pbar ? pbar[k] : zero,
                                                 dp.d_b);
        sdk::copy(s, m, yh, v);
        ctx.trsv(s, 'N', m, d_L, m, v);
        ctx.trsv(s, 'T', m, d_L, m, v);
        for (int sweep = 0; sweep < 2; sweep++) {
            sdk::copy(s, m, yh, r);
            sdk::gemv_n(s, m, m, mk(-1.0), d_B, m, v, one, r);
            ctx.trsv(s, 'N', m, d_L, m, r);
            ctx.trsv(s, 'T', m, d_L, m, r);
            sdk::axpy(s, m, one, r, v);
        }
        plus_A(s, dp, v, d_V + (idx_t)k * nn);
    }
    // G_k = sym(X A^T(v_k) Z^{-1}) - sym(W_k)
    for (int l = 0; l < nB; l++) {
        const idx_t n = dp.bs[l], o = dp.boff[l];
        sdk::gemm(s, 'N', n, n, n, one, d_X + o, n, d_V + o, n, zero, d_P + o, n, K, 0, nn, nn);
        sdk::gemm(s, 'N', n, n, n, one, d_P + o, n, d_Zinv + o, n, zero, d_S + o, n, K, nn, 0, nn);
    }
    k_sym_diff_batch<<<eg, 256, 0, s>>>(nn, nB, dp.d_boff, dp.d_bsize, d_S, d_W, d_S);
    for (int k = 0; k < K; k++) {
        const dd *G = d_S + (idx_t)k * nn, *v = d_v + (idx_t)k * m;
        if (bbar) {
            sdk::copy(s, m, v, bbar + (idx_t)k * m);
            if (pbar) k_add2<<<(m + 255) / 256, 256, 0, s>>>(m, bbar + (idx_t)k * m, nullptr, pbar[k], d_y);
        }
        if (Cbar) sdk::copy(s, nn, G, Cbar + (idx_t)k * nn);
        if (Abar && ntask > 0)
            k_abar<<<ntask, 256, 0, s>>>(ntask, d_tcon, dp.d_task_type, dp.d_task_block, dp.d_task_nnz, dp.d_task_off,
                                         d_aoff, dp.d_sp_row, dp.d_sp_col, dp.d_boff, dp.d_bsize, d_y, v, G, d_X,
                                         Abar + (idx_t)k * na);
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        err = "adjoint: device error";
        return -3;
    }
    return 0;
}

}  // namespace sdp
