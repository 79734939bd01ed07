// Interior-point main loop over device-resident iterates.
//
// The scalar control (phase test, step lengths, centering parameter,
// complementarity, residual ratios) is the reference's host logic evaluated
// in dd on the CPU; every n x n and m x m operation runs on the GPU.  Only
// scalars and block diagonals cross the PCIe boundary during the loop.
#include <atomic>
#include <chrono>
#include <cmath>
#include <deque>
#include <thread>
#include <cstdio>

#include "context.hpp"
#include "solver.hpp"
#include "smallblk.hpp"
#include "steqr.hpp"

namespace sdp {

using sdk::idx_t;

int Context::init(int dev) {
    device = dev;
    if (hipSetDevice(dev) != hipSuccess) return -1;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return -1;
    if (hipMalloc((void **)&d_dot_work, sizeof(dd) * 8192) != hipSuccess) return -1;
    if (hipMalloc((void **)&d_counter, sizeof(unsigned) * 4) != hipSuccess) return -1;
    if (hipMalloc((void **)&d_scal, sizeof(dd) * 64) != hipSuccess) return -1;
    if (hipMalloc((void **)&d_info, sizeof(int64_t)) != hipSuccess) return -1;
    if (hipMalloc((void **)&d_rinv, sizeof(dd) * sdk::POTRF_SCRATCH_DD) != hipSuccess) return -1;
    if (hipMalloc((void **)&d_max_work, sizeof(dd) * sdk::MAX_ABS_BLOCKS) != hipSuccess) return -1;
    if (hipMemset(d_counter, 0, sizeof(unsigned) * 4) != hipSuccess) return -1;
    return 0;
}
void *Context::workspace(size_t bytes) {
    if (bytes > ws_cap) {
        // the old buffer may still be read by work queued on the stream
        if (d_ws) {
            if (hipStreamSynchronize(stream) != hipSuccess) return nullptr;
            (void)hipFree(d_ws);
        }
        d_ws = nullptr;
        ws_cap = 0;
        if (hipMalloc(&d_ws, bytes) != hipSuccess) return nullptr;
        ws_cap = bytes;
    }
    return d_ws;
}
void Context::trsv(hipStream_t s, char trans, sdk::idx_t n, const dd *L, sdk::idx_t lda, dd *x) {
    if (n >= 512) {
        const size_t need = sdk::trsv_coop_bytes(n);
        bool ok = true;
        if (need > tsw_cap) {
            if (d_tsw) (void)hipFree(d_tsw);
            d_tsw = nullptr;
            tsw_cap = 0;
            ok = hipMalloc(&d_tsw, need) == hipSuccess;
            if (ok) tsw_cap = need;
        }
        if (ok && sdk::trsv_lower_coop(s, trans, n, L, lda, x, d_tsw) == 0) return;
    }
    sdk::trsv_lower(s, trans, n, L, lda, x);
}
int Context::binv_prepare(hipStream_t s, sdk::idx_t n, const dd *L, sdk::idx_t lda) {
    const size_t e = sdk::binv_ops_elems(n), need_ops = 3 * e;
    if (need_ops > binv_cap) {
        if (d_binv) (void)hipFree(d_binv);
        d_binv = nullptr;
        binv_cap = 0;
        if (hipMalloc((void **)&d_binv, sizeof(dd) * need_ops) != hipSuccess) return -1;
        binv_cap = need_ops;
    }
    const size_t need = sdk::trsv_coop_bytes(n);
    if (need > tsw_cap) {
        if (d_tsw) (void)hipFree(d_tsw);
        d_tsw = nullptr;
        tsw_cap = 0;
        if (hipMalloc(&d_tsw, need) != hipSuccess) return -1;
        tsw_cap = need;
    }
    // The transposed copy is an optimisation: without it the L^T x solve
    // reads L along rows.
    const size_t need_lt = (size_t)n * n;
    if (need_lt > binv_lt_cap) {
        if (d_binv_lt) (void)hipFree(d_binv_lt);
        d_binv_lt = nullptr;
        binv_lt_cap = 0;
        if (hipMalloc((void **)&d_binv_lt, sizeof(dd) * need_lt) == hipSuccess)
            binv_lt_cap = need_lt;
        else
            (void)hipGetLastError();
    }
    binv_n = n;
    sdk::trsv_binv_prepare(s, n, L, lda, d_binv, d_binv + e, d_binv + 2 * e, d_binv_lt);
    return 0;
}
void Context::trsv_binv(hipStream_t s, char trans, sdk::idx_t n, const dd *L, sdk::idx_t lda, dd *x) {
    const size_t e = sdk::binv_ops_elems(n);
    const bool tr = (trans == 'T' || trans == 't');
    if (n == binv_n && sdk::trsv_binv(s, trans, n, L, lda, x, d_binv + (tr ? e : 0), d_tsw, d_binv_lt) == 0) return;
    sdk::trsv_lower(s, trans, n, L, lda, x);
}
void Context::destroy() {
    if (d_dot_work) (void)hipFree(d_dot_work);
    if (d_counter) (void)hipFree(d_counter);
    if (d_scal) (void)hipFree(d_scal);
    if (d_info) (void)hipFree(d_info);
    if (d_rinv) (void)hipFree(d_rinv);
    if (d_max_work) (void)hipFree(d_max_work);
    ozw.release();
    if (d_tsw) (void)hipFree(d_tsw);
    if (d_ws) (void)hipFree(d_ws);
    d_ws = nullptr;
    ws_cap = 0;
    if (d_binv) (void)hipFree(d_binv);
    d_binv = nullptr;
    binv_cap = 0;
    if (d_binv_lt) (void)hipFree(d_binv_lt);
    d_binv_lt = nullptr;
    binv_lt_cap = 0;
    d_tsw = nullptr;
    tsw_cap = 0;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
}

namespace {
// Lal::getMinEigen (sdpa_linear.cpp:27-122): Lanczos on L M L^T with a start
// vector of ones.  Chains for different (block, X/Z) pairs are independent,
// so they run lock-stepped on their own streams with one host
// synchronisation per step for all of them.  The (k+1)-tridiagonal
// eigenproblem is solved on the host, forming only the last row of its
// eigenvector matrix (the one entry the estimate reads).
static constexpr int LZ_INFLIGHT = 3;  // batches of steps a chain keeps on the device
struct Chain {
    int n = 0, cap = 0;
    const dd *L = nullptr, *M = nullptr;
    dd *r = nullptr, *q = nullptr, *qold = nullptr, *w = nullptr, *tmp = nullptr;
    dd *scal = nullptr;        // device: [alpha, r.r]
    dd *dot_work = nullptr;    // device scalars (beta, 1/beta at [2], [3]) and the step history
    dd *hist = nullptr;        // device: {alpha, r.r} per step
    dd *hh = nullptr;          // pinned host copy of hist
    unsigned *counter = nullptr;
    // batches of steps on the device, oldest first; a batch ends at the
    // step that runs the tridiagonal eigen solve
    struct Batch {
        int first, cnt;
        hipEvent_t ev;
    };
    std::deque<Batch> inflight;
    int ek = 0, ekk = 0;        // the next step to enqueue and its kk
    hipEvent_t evs[LZ_INFLIGHT + 1] = {};  // ring of completion events (in flight + the last one used)
    int evi = 0;
    hipEvent_t last_used = nullptr;
    hipStream_t s = nullptr;
    dd *h = nullptr;           // pinned host copy of scal
    // host state
    ddr alpha, beta, value, min = 1.0e+51, min_old = 1.0e+52, error = 1.0e+10;
    int k = 0, kk = 0;
    bool done = false;
    bool fast = false;  // the Ozaki form: mat-vecs in split order
    std::vector<ddr> diagVec, diagVec2, out, bv, work;
    ddr result;
};

static constexpr int HIST_CAP = 2040;  // steps of one run (at most sqrt(n) + 11)

struct ChainPool {
    std::vector<Chain> chains;
    std::vector<void *> dmem, hmem;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> events;
    ~ChainPool() {
        drain();
        for (void *p : dmem) (void)hipFree(p);
        for (void *p : hmem) (void)hipHostFree(p);
        for (hipStream_t s : streams) (void)hipStreamDestroy(s);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
    // waits for the steps a chain ran past its end (they still read L and M)
    void drain() {
        for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
    }
    // make sure chains [0, cnt) exist with vectors of length >= n
    bool ensure(int cnt, int n) {
        while ((int)chains.size() < cnt) {
            Chain c;
            dd *v = nullptr;
            if (hipMalloc((void **)&v, sizeof(dd) * (5 * (size_t)n + 2 + 4 + 4096)) != hipSuccess) return false;
            dmem.push_back(v);
            c.r = v;
            c.q = v + n;
            c.qold = v + 2 * (size_t)n;
            c.w = v + 3 * (size_t)n;
            c.tmp = v + 4 * (size_t)n;
            c.scal = v + 5 * (size_t)n;
            c.dot_work = v + 5 * (size_t)n + 2;
            if (hipMalloc((void **)&c.counter, sizeof(unsigned) * 2) != hipSuccess) return false;
            dmem.push_back(c.counter);
            (void)hipMemset(c.counter, 0, sizeof(unsigned) * 2);
            if (hipHostMalloc((void **)&c.h, sizeof(dd) * 2) != hipSuccess) return false;
            hmem.push_back(c.h);
            c.hist = c.dot_work + 8;
            if (hipHostMalloc((void **)&c.hh, sizeof(dd) * 2 * HIST_CAP) != hipSuccess) return false;
            hmem.push_back(c.hh);
            for (hipEvent_t &e : c.evs) {
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return false;
                events.push_back(e);
            }
            hipStream_t s;
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return false;
            streams.push_back(s);
            c.s = s;
            c.n = c.cap = n;
            chains.push_back(c);
        }
        for (int i = 0; i < cnt; i++)
            if (chains[i].cap < n) return false;
        return true;
    }
};

bool chain_active(Chain &c) {
    const int nDim = c.n;
    return c.k < nDim && c.k < sqrt((ddr)nDim) + 10 && c.beta > 1.0e-16 &&
           (abs(c.min - c.min_old) > (1.0e-5) * abs(c.min) + (1.0e-8) ||
            abs(c.error * c.beta) > (1.0e-2) * abs(c.min) + (1.0e-4));
}

// Enqueue steps first .. first + cnt - 1: each forms q from the device
// scalars, the three mat-vecs, then alpha, r and r.r into hist[first + j]
// and the next scalars (Lal::getMinEigen's loop body, sdpa_linear.cpp:64-82),
// then the copy of their history to the host and a completion event.
void chain_enqueue(Chain &c, int T, int first, int cnt) {
    hipStream_t s = c.s;
    const int nDim = c.n;
    dd *sc = c.dot_work;
    for (int j = 0; j < cnt; j++) {
        // the Ozaki form: the same step in three launches
        if (c.fast && sdk::lanczos_step_fast(s, nDim, c.L, c.M, c.r, c.q, c.qold, c.w, c.tmp, sc,
                                             c.hist + 2 * (size_t)(first + j), c.counter))
            continue;
        sdk::lanczos_q_dev(s, nDim, sc, c.r, c.q, c.qold);  // qold = q; q = r / beta
        if (c.fast) {
            sdk::trmv_lower_fast(s, 'T', nDim, c.L, nDim, c.q, c.w);
            sdk::gemv_n_fast(s, nDim, nDim, mk(1.0), c.M, nDim, c.w, mk(0.0), c.tmp);
            sdk::trmv_lower_fast(s, 'N', nDim, c.L, nDim, c.tmp, c.w);
        } else {
            sdk::trmv_lower(s, 'T', nDim, c.L, nDim, c.q, c.w);
            sdk::gemv_n(s, nDim, nDim, mk(1.0), c.M, nDim, c.w, mk(0.0), c.tmp);
            sdk::trmv_lower(s, 'N', nDim, c.L, nDim, c.tmp, c.w);
        }
        // alpha = q.w; r = w - alpha q - beta qold; r.r; next beta, 1/beta
        if (c.fast)
            sdk::lanczos_tail_fast(s, nDim, c.q, c.qold, c.w, c.r, sc, c.hist + 2 * (size_t)(first + j));
        else
            sdk::lanczos_tail_dev(s, nDim, T, c.q, c.qold, c.w, c.r, sc, c.hist + 2 * (size_t)(first + j));
    }
    (void)hipMemcpyAsync(c.hh + 2 * (size_t)first, c.hist + 2 * (size_t)first, sizeof(dd) * 2 * cnt,
                         hipMemcpyDeviceToHost, s);
    hipEvent_t ev = c.evs[c.evi];
    c.evi = (c.evi + 1) % (LZ_INFLIGHT + 1);
    (void)hipEventRecord(ev, s);
    c.inflight.push_back({first, cnt, ev});
}

// Steps from k on (kk steps since the last eigen solve) up to and including
// the next one that runs the tridiagonal eigen solve (whose outcome may end
// the chain).  The schedule depends only on k and kk.
int chain_batch(const Chain &c, int k, int kk) {
    const int nDim = c.n;
    int cnt = 0;
    while (k < nDim && k < sqrt((ddr)nDim) + 10 && k < HIST_CAP) {
        cnt++;
        if (kk >= sqrt((ddr)k) || k == nDim - 1 || k > sqrt((ddr)nDim + 9)) break;
        k++;
        kk++;
    }
    return cnt;
}

void chain_finish_step(Chain &c, const dd *rec) {
    const int nDim = c.n, k = c.k;
    c.alpha = ddr(rec[0]);
    c.diagVec[k] = c.alpha;
    if (c.kk >= sqrt((ddr)k) || k == nDim - 1 || k > sqrt((ddr)nDim + 9)) {
        c.kk = 0;
        c.out = c.diagVec;
        c.bv = c.diagVec2;
        c.out[nDim - 1] = c.diagVec[k];
        c.bv[nDim - 1] = 0.0;
        const int kp1 = k + 1;
        std::vector<ddr> z((size_t)kp1 * kp1);
        int64_t info = steqr_rows(kp1, c.out.data(), c.bv.data(), z.data(), kp1, c.work.data(), k, k + 1);
        if (info > 0) {  // "cannot converge": leave the loop with beta unchanged
            c.done = true;
            return;
        }
        c.min_old = c.min;
        c.min = c.out[0];
        c.error = z[k];
    }
    c.value = ddr(rec[1]);
    c.beta = sqrt(c.value);
    c.diagVec2[k] = c.beta;
    ++c.k;
    ++c.kk;
}

// Runs chains[0..cnt) to completion; chain i computes lambda_min(L_i M_i L_i^T).
void run_chains(ChainPool &pool, int cnt, int T, hipStream_t parent, bool fast) {
    hipEvent_t ready;
    (void)hipEventCreateWithFlags(&ready, hipEventDisableTiming);
    (void)hipEventRecord(ready, parent);
    for (int i = 0; i < cnt; i++) {
        Chain &c = pool.chains[i];
        (void)hipStreamWaitEvent(c.s, ready, 0);
        c.fast = fast;
        const int nDim = c.n;
        c.min = 1.0e+51;
        c.min_old = 1.0e+52;
        c.error = 1.0e+10;
        c.k = c.kk = 0;
        c.done = false;
        c.diagVec.assign(nDim, ddr(1.0e+50));
        c.diagVec2.assign(nDim, ddr(0.0));
        c.out.assign(nDim, ddr(0.0));
        c.bv.assign(nDim, ddr(0.0));
        c.work.assign(2 * (size_t)nDim + 2, ddr(0.0));
        sdk::fill(c.s, nDim, mk(0.0), c.q);
        sdk::fill(c.s, nDim, mk(1.0), c.r);
        c.beta = sqrt((ddr)nDim);
        // device scalars of the first step: beta and 1/beta
        c.h[0] = c.beta.v;
        c.h[1] = (ddr(1.0) / c.beta).v;
        (void)hipMemcpyAsync(c.dot_work + 2, c.h, sizeof(dd) * 2, hipMemcpyHostToDevice, c.s);
    }
    // Each chain keeps its next batch running while the host processes the
    // previous batch's eigen solve: the batch schedule depends only on the
    // step counters and the device forms its own scalars, so only the
    // decision to stop waits for the host.  Steps past that decision are
    // dropped (their results are never read, and nothing waits for them).
    auto enqueue_next = [&](Chain &c) {
        const int n_steps = chain_batch(c, c.ek, c.ekk);
        if (n_steps == 0) return false;
        chain_enqueue(c, T, c.ek, n_steps);
        c.ek += n_steps;
        c.ekk = 1;  // a batch ends at an eigen solve, after which kk restarts
        return true;
    };
    for (int i = 0; i < cnt; i++) {
        Chain &c = pool.chains[i];
        c.inflight.clear();
        c.ek = c.ekk = 0;
        c.last_used = nullptr;
        for (int b = 0; b < LZ_INFLIGHT && enqueue_next(c); b++) {
        }
    }
    // The host work of a chain (reading its batches, the tridiagonal eigen
    // solves, enqueueing) runs on one of up to LZ_THREADS host threads, the
    // chains dealt round-robin; each thread visits its chains in turn.
    auto drive = [&](int i0, int stride) {
        for (;;) {
            bool any = false;
            for (int i = i0; i < cnt; i += stride) {
                Chain &c = pool.chains[i];
                if (c.done || c.inflight.empty()) continue;
                any = true;
                const Chain::Batch b = c.inflight.front();
                c.inflight.pop_front();
                (void)hipEventSynchronize(b.ev);
                c.last_used = b.ev;
                for (int j = 0; j < b.cnt; j++) {
                    // within a batch the convergence test cannot change (min
                    // and error move only at the eigen solve that closes the
                    // batch); beta can end the chain early, and the later
                    // steps the device already ran are then dropped
                    if (!chain_active(c)) {
                        c.done = true;
                        break;
                    }
                    chain_finish_step(c, c.hh + 2 * (size_t)(b.first + j));
                    if (c.done) break;
                }
                if (!c.done && c.inflight.size() < LZ_INFLIGHT) enqueue_next(c);
            }
            if (!any) break;
        }
    };
    constexpr int LZ_THREADS = 4;
    const int nth = cnt < LZ_THREADS ? cnt : LZ_THREADS;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::vector<std::thread> helpers;
    for (int t = 1; t < nth; t++)
        helpers.emplace_back([&, t] {
            (void)hipSetDevice(dev);  // the current device is per thread
            drive(t, nth);
        });
    drive(0, nth);
    for (std::thread &h : helpers) h.join();
    for (int i = 0; i < cnt; i++) {
        Chain &c = pool.chains[i];
        c.result = c.min - abs(c.error * c.beta);
        if (c.last_used) (void)hipStreamWaitEvent(parent, c.last_used, 0);
    }
    (void)hipEventDestroy(ready);
}

}  // namespace

namespace {

struct DevBuf {
    dd *p = nullptr;
    idx_t n = 0;
};

class Solver {
  public:
    Solver(Context &c, const HostProblem &h, DeviceProblem &d, const Params &p, const InitPoint *ip,
           const sps::Device *sp)
        : ctx(c), hp(h), dp(d), prm(p), init(ip), spd(sp) {}
    ~Solver() {
        pool.drain();
        if (s2) (void)hipStreamSynchronize(s2);
        if (s3) (void)hipStreamSynchronize(s3);
        for (void *q : mem) (void)hipFree(q);
        if (h_sc) (void)hipHostFree(h_sc);
        if (h_diag) (void)hipHostFree(h_diag);
        if (h_infos) (void)hipHostFree(h_infos);
        if (h_small_out) (void)hipHostFree(h_small_out);
        if (h_small_info) (void)hipHostFree(h_small_info);
        if (h_lp_eig) (void)hipHostFree(h_lp_eig);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_join3) (void)hipEventDestroy(ev_join3);
        if (la.ready) (void)hipEventDestroy(la.ready);
        if (la.done) (void)hipEventDestroy(la.done);
        for (hipEvent_t e : {ev_t0, ev_t1, ev_t2})
            if (e) (void)hipEventDestroy(e);
        if (s2) (void)hipStreamDestroy(s2);
        if (s3) (void)hipStreamDestroy(s3);
        ozw2.release();
        ozw3.release();
    }
    int run(SolveOutput &out, std::vector<dd> *xo, std::vector<dd> *yo, std::vector<dd> *zo);

  private:
    Context &ctx;
    const HostProblem &hp;
    DeviceProblem &dp;
    const Params &prm;
    const InitPoint *init;
    // the sparse-Schur path (null: dense B): the pattern's values and the
    // solves' work vector
    const sps::Device *spd;
    dd *spv = nullptr, *spx = nullptr;
    dd *initX = nullptr, *initZ = nullptr;  // the initial point's X and Z (kept for SolveInfo)
    sdk::DotTask *d_tasks_init = nullptr;   // (initX, Z) and (X, initZ) per block
    std::vector<void *> mem;
    hipStream_t s = nullptr;
    int m = 0, nB = 0, nDim = 0;
    idx_t nn = 0;
    // LP entries: the last nlp values of every iterate (after the nn SDP
    // values); nt = nn + nlp.  Their operations follow the LP loops of the
    // reference's linear-space routines (different roundings from the SDP
    // blocks' Raxpy forms), so whole-space updates split at nn.
    idx_t nlp = 0, nt = 0;
    int slot_lp = 0;  // first LP-term slot
    std::vector<int> small_blk;  // SDP blocks of size <= 20 (Rsyev step length)
    sdk::SmallEigTask *d_small_tasks = nullptr;
    dd *d_small_out = nullptr, *h_small_out = nullptr;
    int *d_small_info = nullptr, *h_small_info = nullptr;
    dd *d_lp_eig = nullptr, *h_lp_eig = nullptr;
    // iterates and work spaces (concatenated SDP blocks)
    dd *X, *Z, *invCholX, *invCholZ, *invZ, *DxMat, *DzMat, *r_zinv, *x_rd_zinv, *DLS1, *DLS2, *dualMat;
    dd *y, *dy, *g, *primalVec, *bMat;
    // scalar slots: [0, 3 nB) block inner products, then C, b and max-norm
    // results; h_sc is their pinned host mirror.  d_diag gathers the X and Z
    // diagonals (2 nDim) for SolveInfo.
    dd *d_sc = nullptr, *h_sc = nullptr, *d_diag = nullptr, *h_diag = nullptr;
    // scalar slots: [0, 3 nB) direction products per block, [3 nB, 4 nB)
    // X . Z per block, [4 nB, 6 nB) the initial-point products; then named
    // slots; then 6 nlp LP terms (3 direction, 1 X . Z, 2 initial-point)
    int slot_xz = 0, slot_init = 0, slot_C = 0, slot_b = 0, slot_C2 = 0, slot_b2 = 0, slot_max = 0, slot_max2 = 0,
        slot_bmax = 0, nslots = 0;
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr, ev_t2 = nullptr;  // Schur build / Cholesky timing
    // Fixed dense inner-product task lists: (X, Z) per block, and (X, dZ),
    // (dX, Z), (dX, dZ) per block; partial sums of their chunks.
    sdk::DotTask *d_tasks_xz = nullptr, *d_tasks3 = nullptr;
    dd *d_partial = nullptr;
    idx_t max_blk_elems = 0;
    dd *gsplit = nullptr;  // the Ozaki form's split products: GEMM_KS parts of the largest block
    // second stream for the Z factorisation, per-block Cholesky info slots
    // X's factorisation runs on s3, Z's on s2, so that the main stream's
    // residual and inner-product work of the same step overlaps both
    // (the process has few hardware queues: s2 doubles as the look-ahead
    // helper of the Schur Cholesky, which runs at another time)
    hipStream_t s2 = nullptr, s3 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join3 = nullptr;
    sdk::OzLookAhead la;
    int64_t *d_infos = nullptr, *h_infos = nullptr;
    dd *d_rinv2 = nullptr, *d_rinv3 = nullptr;
    oz::Work ozw2, ozw3;      // Ozaki planes for the second and third streams
    dd *bSave = nullptr, *rvec = nullptr, *dyTry = nullptr;  // Ozaki form: unfactored B, residual
    int refine_accepted = 0;
    ddr bmax = 0.0;
    double t_schur = 0, t_chol = 0, t_solve = 0, t_lanczos = 0, t_inverse = 0;

    dd *alloc(idx_t n) {
        dd *p = nullptr;
        if (n <= 0) n = 1;
        if (hipMalloc((void **)&p, sizeof(dd) * n) != hipSuccess) return nullptr;
        mem.push_back(p);
        return p;
    }
    dd *blk(dd *base, int l) { return base + dp.boff[l]; }
    int bsz(int l) const { return hp.bs[l]; }
    void sync() { (void)hipStreamSynchronize(s); }
    static double now() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
    void dot(idx_t n, const dd *x, const dd *y, int slot) {
        if (ctx.path >= 1)
            sdk::dot_fast(s, n, x, y, d_sc + slot, ctx.d_dot_work, ctx.d_counter + 2);
        else
            sdk::dot(s, n, x, y, d_sc + slot, ctx.dot_T, ctx.d_dot_work, ctx.d_counter);
    }
    // Download scalar slots [first, first + cnt) into h_sc.
    const dd *get_scalars(int first, int cnt) {
        (void)hipMemcpyAsync(h_sc + first, d_sc + first, sizeof(dd) * cnt, hipMemcpyDeviceToHost, s);
        sync();
        return h_sc + first;
    }
    // Scalar batches: enqueue several reductions, then one download and one
    // synchronisation for all of them (fetch_all) -- or a download enqueued
    // ahead of a later synchronisation (prefetch_all).
    void prefetch_all() {
        (void)hipMemcpyAsync(h_sc, d_sc, sizeof(dd) * nslots, hipMemcpyDeviceToHost, s);
    }
    void fetch_all() {
        prefetch_all();
        sync();
    }
    ddr sc(int slot) const { return ddr(h_sc[slot]); }
    void enq_inner_C(const dd *W, int slot) { sdp::inner_C(s, dp, W, d_sc + slot, ctx.dot_T, ctx.path >= 1); }
    void enq_inner_b(const dd *v, int slot) { dot(m, dp.d_b, v, slot); }
    void enq_max(const dd *x, idx_t len, int slot) {
        sdk::max_abs(s, len, x, d_sc + slot, ctx.d_max_work, ctx.d_counter + 1);
    }
    // X . dZ, dX . Z, dX . dZ (block dots, then LP terms)
    void enq_dirs() {
        if (nB) dots(d_tasks3, 3 * nB);
        if (nlp) {
            sdk::lp_prod(s, nlp, X + nn, DzMat + nn, d_sc + slot_lp);
            sdk::lp_prod(s, nlp, DxMat + nn, Z + nn, d_sc + slot_lp + nlp);
            sdk::lp_prod(s, nlp, DxMat + nn, DzMat + nn, d_sc + slot_lp + 2 * nlp);
        }
    }
    void read_dirs(ddr &r0, ddr &r1, ddr &r2) const {
        ddr r[3];
        for (int q = 0; q < 3; q++) {
            r[q] = 0.0;
            for (int l = 0; l < nB; l++) r[q] += sc(q * nB + l);
            for (idx_t l = 0; l < nlp; l++) r[q] += sc(slot_lp + q * (int)nlp + (int)l);
        }
        r0 = r[0];
        r1 = r[1];
        r2 = r[2];
    }
    void enq_xz() {
        if (nB) dots(d_tasks_xz, nB);
        if (nlp) sdk::lp_prod(s, nlp, X + nn, Z + nn, d_sc + slot_lp + 3 * nlp);
    }
    ddr read_xz() const {
        ddr r = 0.0;
        for (int l = 0; l < nB; l++) r += sc(slot_xz + l);
        for (idx_t l = 0; l < nlp; l++) r += sc(slot_lp + 3 * (int)nlp + (int)l);
        return r;
    }
    void enq_init() {
        if (nB) dots(d_tasks_init, 2 * nB);
        if (nlp) {
            sdk::lp_prod(s, nlp, initX + nn, Z + nn, d_sc + slot_lp + 4 * nlp);
            sdk::lp_prod(s, nlp, X + nn, initZ + nn, d_sc + slot_lp + 5 * nlp);
        }
    }
    void read_init(ddr &x0z, ddr &xz0) const {
        ddr r[2];
        for (int q = 0; q < 2; q++) {
            r[q] = 0.0;
            for (int l = 0; l < nB; l++) r[q] += sc(slot_init + q * nB + l);
            for (idx_t l = 0; l < nlp; l++) r[q] += sc(slot_lp + (4 + q) * (int)nlp + (int)l);
        }
        x0z = r[0];
        xz0 = r[1];
    }

    // ---- dense linear-space helpers (sdpa_linear.cpp Lal::*) ----
    // DenseLinearSpace inner product X . Z: per-block Rdot (all blocks in one
    // launch), summed in block order.
    // the solver's batched dots: the reference's partition order, or the
    // Ozaki form's fixed tree (d_partial holds 3 nB T entries)
    void dots(const sdk::DotTask *tasks, int cnt) {
        if (ctx.path >= 1)
            sdk::dot_tasks_fast(s, tasks, cnt, max_blk_elems, 3 * nB * ctx.dot_T / cnt, d_partial);
        else
            sdk::dot_tasks(s, tasks, cnt, max_blk_elems, ctx.dot_T, d_partial);
    }
    // The LP terms a_l b_l follow the block sums in ascending l
    // (Lal::getInnerProduct, sdpa_linear.cpp:1603-1611).
    ddr inner_xz() {
        enq_xz();
        fetch_all();
        return read_xz();
    }
    ddr inner_C(const dd *W) {
        sdp::inner_C(s, dp, W, d_sc + slot_C, ctx.dot_T, ctx.path >= 1);
        return ddr(*get_scalars(slot_C, 1));
    }
    ddr inner_b(const dd *v) {
        dot(m, dp.d_b, v, slot_b);
        return ddr(*get_scalars(slot_b, 1));
    }
    ddr max_abs(const dd *x, idx_t len) {
        sdk::max_abs(s, len, x, d_sc + slot_max, ctx.d_max_work, ctx.d_counter + 1);
        return ddr(*get_scalars(slot_max, 1));
    }
    // ret = a*b*c per block (Jal::ns_jordan_triple_product)
    void triple(dd *ret, const dd *a, const dd *b, const dd *c, dd *work) {
        if (ctx.path == 2) {
            for (int l = 0; l < nB; l++) {
                int n = bsz(l);
                idx_t o = dp.boff[l];
                oz_check(oz::gemm(s, ctx.ozw, 'N', 'N', n, n, n, 1, a + o, n, b + o, n, 0, work + o, n, false, ctx.oz_slices));
                oz_check(oz::gemm(s, ctx.ozw, 'N', 'N', n, n, n, 1, work + o, n, c + o, n, 0, ret + o, n, false, ctx.oz_slices));
            }
            if (nlp) sdk::lp_triple(s, nlp, a + nn, b + nn, c + nn, ret + nn);
            return;
        }
        // the Ozaki form's products in four k parts (shorter dd chains)
        for (int l = 0; l < nB; l++) {
            int n = bsz(l);
            idx_t o = dp.boff[l];
            if (gsplit)
                sdk::gemm_split(s, n, n, n, a + o, n, b + o, n, work + o, n, gsplit);
            else
                sdk::gemm(s, 'N', n, n, n, mk(1.0), a + o, n, b + o, n, mk(0.0), work + o, n);
        }
        for (int l = 0; l < nB; l++) {
            int n = bsz(l);
            idx_t o = dp.boff[l];
            if (gsplit)
                sdk::gemm_split(s, n, n, n, work + o, n, c + o, n, ret + o, n, gsplit);
            else
                sdk::gemm(s, 'N', n, n, n, mk(1.0), work + o, n, c + o, n, mk(0.0), ret + o, n);
        }
        if (nlp) sdk::lp_triple(s, nlp, a + nn, b + nn, c + nn, ret + nn);
    }
    std::atomic<bool> oz_failed{false};  // set from the X/Z enqueue threads too
    void oz_check(int rc) {
        if (rc != 0) oz_failed = true;
    }
    // Cholesky of the Schur matrix or of an X/Z block in the current form.
    void potrf(hipStream_t st, idx_t n, dd *A, int64_t *info, dd *rinv, oz::Work &w) {
        if (ctx.path >= 1)
            oz_check(sdk::potrf_lower_oz(st, n, A, n, info, rinv, w, ctx.oz_slices, st == s ? &la : nullptr));
        else
            sdk::potrf_lower(st, n, A, n, info, rinv);
    }
    // dy = B^{-1} g with the factor L in bMat.  In the Ozaki form this is
    // followed by up to refine_steps sweeps of dd iterative refinement
    // against the saved B: r = g - B dy, c = L^-T L^-1 r, dy += c.  A sweep
    // runs only while max|r| is above the dd rounding floor 64 m 2^-104
    // max|B| max|dy| (below it a correction only moves dy within rounding
    // noise, which changes the iterates without improving them) and is kept
    // only if it at least halves max|r|.
    // The Schur solves do not wait for the device.
    bool binv = false;
    void trsv_chk(char trans, idx_t n, const dd *L, idx_t lda, dd *x) {
        if (binv)
            ctx.trsv_binv(s, trans, n, L, lda, x);
        else
            ctx.trsv(s, trans, n, L, lda, x);
    }
    void residual(const dd *x, dd *r) {
        sdk::copy(s, m, g, r);
        sdk::gemv_n_fast(s, m, m, mk(-1.0), bSave, m, x, mk(1.0), r);
    }
    void schur_solve() {
        if (spd) {
            // compute_DyVec's sparse branch (sdpa_newton.cpp:1298-1302)
            sps::solve(s, *spd, spv, g, dy, spx);
            return;
        }
        sdk::copy(s, m, g, dy);
        trsv_chk('N', m, bMat, m, dy);
        trsv_chk('T', m, bMat, m, dy);
        if (ctx.path == 0 || ctx.refine_steps == 0) return;
        // max|r| and max|dy| of each candidate come down together
        residual(dy, rvec);
        enq_max(rvec, m, slot_max);
        enq_max(dy, m, slot_max2);
        const dd *h = get_scalars(slot_max, 2);
        ddr rn = h[0], mdy = h[1];
        const ddr floor_scale = 64.0 * m * ldexp(1.0, -104) * bmax;
        for (int it = 0; it < ctx.refine_steps; it++) {
            if (!(rn > floor_scale * mdy)) break;
            sdk::copy(s, m, dy, dyTry);
            trsv_chk('N', m, bMat, m, rvec);
            trsv_chk('T', m, bMat, m, rvec);
            sdk::axpy(s, m, mk(1.0), rvec, dyTry);
            residual(dyTry, rvec);
            enq_max(rvec, m, slot_max);
            enq_max(dyTry, m, slot_max2);
            h = get_scalars(slot_max, 2);
            ddr rt = h[0], mtry = h[1];
            if (!(rt * 2.0 <= rn)) break;
            sdk::copy(s, m, dyTry, dy);
            rn = rt;
            mdy = mtry;
            refine_accepted++;
        }
    }
    // ret = a + s*b (Lal::plus) over the whole space
    void plus(dd *ret, const dd *a, const dd *b, dd sc) {
        sdk::copy(s, nn, a, ret);
        sdk::axpy(s, nn, sc, b, ret);
        if (nlp) sdk::lp_plus(s, nlp, a + nn, b + nn, sc, ret + nn);
    }
    // a += s*b in place: the SDP blocks by Raxpy, the LP entries as Lal::plus
    void plus_in(dd *a, const dd *b, dd sc) {
        sdk::axpy(s, nn, sc, b, a);
        if (nlp) sdk::lp_plus(s, nlp, a + nn, b + nn, sc, a + nn);
    }

    // ---- pieces of the algorithm ----
    // Residuals::compute (sdpa_dataset.cpp:625-645); the norms go to
    // slot_max (primal) and slot_max2 (dual)
    void enq_residuals() {
        sdp::inner_A(s, dp, X, primalVec, ctx.dot_T, ctx.path >= 1);
        sdk::scal(s, m, mk(-1.0), primalVec);
        sdk::axpy(s, m, mk(1.0), dp.d_b, primalVec);
        sdp::plus_A(s, dp, y, dualMat);
        sdk::scal(s, nt, mk(-1.0), dualMat);
        sdp::plus_C(s, dp, dualMat);
        plus_in(dualMat, Z, mk(-1.0));
        enq_max(primalVec, m, slot_max);
        enq_max(dualMat, nt, slot_max2);
    }
    void residuals(ddr &normP, ddr &normD) {
        enq_residuals();
        const dd *h = get_scalars(slot_max, 2);
        normP = h[0];
        normD = h[1];
    }
    // Jal::getInvChol / getInvCholAndInv (sdpa_jordan.cpp:123-170) for every
    // block of A, enqueued on `st` without synchronisation; block l's
    // Cholesky info goes to infos[l].
    void inv_chol_async(hipStream_t st, dd *A, dd *invChol, dd *work, dd *inv, int64_t *infos, int64_t *lp_info,
                        dd *rinv) {
        for (int l = 0; l < nB; l++) {
            int n = bsz(l);
            idx_t o = dp.boff[l];
            sdk::copy(st, (idx_t)n * n, A + o, work + o);
            oz::Work &w = st == s ? ctx.ozw : (st == s2 ? ozw2 : ozw3);
            potrf(st, n, work + o, infos + l, rinv, w);
            sdk::zero_upper(st, n, work + o);
            // path 2 (dense iterates): the inverse factor's updates and
            // L^-T L^-1 as Ozaki products, like the direction's GEMMs
            if (ctx.path == 2)
                oz_check(sdk::inv_lower_oz(st, n, work + o, n, invChol + o, n, w, ctx.oz_slices));
            else
                sdk::inv_lower(st, n, work + o, n, invChol + o, n, ctx.path >= 1);
        }
        if (inv)
            for (int l = 0; l < nB; l++) {
                int n = bsz(l);
                idx_t o = dp.boff[l];
                if (ctx.path == 2)
                    oz_check(oz::gemm(st, st == s ? ctx.ozw : (st == s2 ? ozw2 : ozw3), 'T', 'N', n, n, n, 1, invChol + o, n, invChol + o,
                                      n, 0, inv + o, n, false, ctx.oz_slices));
                else
                    sdk::trmm_lltn_self(st, n, invChol + o, n, inv + o, n);
            }
        if (nlp) sdk::lp_inv_chol(st, nlp, A + nn, invChol + nn, inv ? inv + nn : nullptr, lp_info);
    }
    // Solutions::update's factorisations: X on the main stream and Z on the
    // second one, concurrently.  Returns false when either Cholesky fails.
    // (The results after a failed block are never used: the loop ends.)
    // X's factorisation on the main stream, Z's on the second one; work
    // enqueued on the main stream between start and join overlaps Z's.
    void enq_inv_chol_start() {
        enq_inv_chol_fork();
        enq_inv_chol_launch();
    }
    // the fork point (the factorisations wait for the work on s so far) ...
    void enq_inv_chol_fork() {
        (void)hipEventRecord(ev_fork, s);
        (void)hipStreamWaitEvent(s2, ev_fork, 0);
        (void)hipStreamWaitEvent(s3, ev_fork, 0);
    }
    // ... and their launches, which may be enqueued after later work on s
    // (their many small launches then do not hold that work back)
    // X's launches and Z's are enqueued from two host threads: enqueued in
    // turn, Z's first launch waited for all of X's (hundreds of launches at
    // n = 800), so the two factorisations barely overlapped on the device
    void enq_inv_chol_launch() {
        int dev = 0;
        (void)hipGetDevice(&dev);
        std::thread tz([&] {
            (void)hipSetDevice(dev);  // the current device is per thread
            inv_chol_async(s2, Z, invCholZ, DLS2, invZ, d_infos + nB, d_infos + 2 * nB + 1, d_rinv2);
        });
        inv_chol_async(s3, X, invCholX, DLS1, nullptr, d_infos, d_infos + 2 * nB, d_rinv3);
        tz.join();
    }
    void enq_inv_chol_join() {
        (void)hipEventRecord(ev_join, s2);
        (void)hipStreamWaitEvent(s, ev_join, 0);
        (void)hipEventRecord(ev_join3, s3);
        (void)hipStreamWaitEvent(s, ev_join3, 0);
        (void)hipMemcpyAsync(h_infos, d_infos, sizeof(int64_t) * (2 * nB + 2), hipMemcpyDeviceToHost, s);
    }
    void enq_inv_chol_xz() {
        enq_inv_chol_start();
        enq_inv_chol_join();
    }
    // a panel workgroup's wait for its diagonal block ran out (an error of
    // the run, not a property of the matrix): the solve returns an error
    bool wait_failed = false;
    bool inv_chol_ok() {  // after a synchronisation
        for (int q = 0; q < 2 * nB + (nlp ? 2 : 0); q++) {
            if (q < 2 * nB && h_infos[q] == sdk::POTRF_WAIT_FAILED) wait_failed = true;
            if (h_infos[q] != 0) return false;
        }
        return true;
    }
    bool inv_chol_xz() {
        double t0 = now();
        enq_inv_chol_xz();
        sync();
        t_inverse += now() - t0;
        return inv_chol_ok();
    }
    ChainPool pool;
    // Jal::getMinEigen (sdpa_jordan.cpp:55-88) for (L_X^{-1}, dX) and (L_Z^{-1}, dZ):
    // all blocks of both pairs at once -- Lanczos for blocks of size > 20,
    // Rsyev for the smaller ones, (m l) l for the LP entries.  The minimum
    // over blocks does not depend on their order.  Returns false when an
    // Rsyev did not converge (the reference ends the run there).
    bool min_eigen_xz(ddr &mx, ddr &mz) {
        double t0 = now();
        std::vector<int> big;
        int maxn = 0;
        for (int l = 0; l < nB; l++)
            if (bsz(l) > sdk::SMALL_EIG_MAXN) {
                big.push_back(l);
                maxn = std::max(maxn, bsz(l));
            }
        const int nbig = (int)big.size(), nsmall = (int)small_blk.size();
        mx = 1.0E50;
        mz = 1.0E50;
        if (nsmall) sdk::small_min_eig(s, d_small_tasks, 2 * nsmall, ctx.dot_T, d_small_out, d_small_info);
        if (nlp) {
            sdk::lp_eig(s, nlp, DxMat + nn, invCholX + nn, d_lp_eig);
            sdk::lp_eig(s, nlp, DzMat + nn, invCholZ + nn, d_lp_eig + nlp);
        }
        if (nbig) {
            if (!pool.ensure(2 * nbig, maxn)) {
                ctx.set_error("Lanczos workspace allocation failed");
                mx = mz = ddr(NAN);
                return false;
            }
            for (int q = 0; q < nbig; q++) {
                const int l = big[q];
                Chain &cx = pool.chains[q];
                cx.n = bsz(l);
                cx.L = invCholX + dp.boff[l];
                cx.M = DxMat + dp.boff[l];
                Chain &cz = pool.chains[nbig + q];
                cz.n = bsz(l);
                cz.L = invCholZ + dp.boff[l];
                cz.M = DzMat + dp.boff[l];
            }
            run_chains(pool, 2 * nbig, ctx.dot_T, s, ctx.path >= 1);
            for (int q = 0; q < nbig; q++) {
                if (pool.chains[q].result < mx) mx = pool.chains[q].result;
                if (pool.chains[nbig + q].result < mz) mz = pool.chains[nbig + q].result;
            }
        }
        bool ok = true;
        if (nsmall || nlp) {
            if (nsmall) {
                (void)hipMemcpyAsync(h_small_out, d_small_out, sizeof(dd) * 2 * nsmall, hipMemcpyDeviceToHost, s);
                (void)hipMemcpyAsync(h_small_info, d_small_info, sizeof(int) * 2 * nsmall, hipMemcpyDeviceToHost, s);
            }
            if (nlp) (void)hipMemcpyAsync(h_lp_eig, d_lp_eig, sizeof(dd) * 2 * nlp, hipMemcpyDeviceToHost, s);
            sync();
            for (int q = 0; q < nsmall; q++) {
                if (h_small_info[q] != 0 || h_small_info[nsmall + q] != 0) ok = false;
                if (ddr(h_small_out[q]) < mx) mx = h_small_out[q];
                if (ddr(h_small_out[nsmall + q]) < mz) mz = h_small_out[nsmall + q];
            }
            for (idx_t q = 0; q < nlp; q++) {
                if (ddr(h_lp_eig[q]) < mx) mx = h_lp_eig[q];
                if (ddr(h_lp_eig[nlp + q]) < mz) mz = h_lp_eig[nlp + q];
            }
        } else {
            sync();
        }
        t_lanczos += now() - t0;
        return ok;
    }
    // x0zMat += lambda * Z_ii, xMatz0 += lambda * X_ii, blocks in order:
    // both diagonals gathered on the device, one contiguous download
    void enq_diag() {
        idx_t off = 0;
        for (int l = 0; l < nB; l++) {
            sdk::gather_diag(s, bsz(l), blk(Z, l), d_diag + off);
            sdk::gather_diag(s, bsz(l), blk(X, l), d_diag + nDim + off);
            off += bsz(l);
        }
        if (nlp) {  // the LP entries follow the SDP diagonals (sdpa_parts.cpp:767-770)
            sdk::copy(s, nlp, Z + nn, d_diag + off);
            sdk::copy(s, nlp, X + nn, d_diag + nDim + off);
        }
        (void)hipMemcpyAsync(h_diag, d_diag, sizeof(dd) * 2 * (size_t)nDim, hipMemcpyDeviceToHost, s);
    }
    void read_diag(double lam, ddr &x0z, ddr &xz0) const {
        for (int i = 0; i < nDim; i++) {
            x0z += lam * ddr(h_diag[i]);
            xz0 += lam * ddr(h_diag[nDim + i]);
        }
    }
};

int Solver::run(SolveOutput &res, std::vector<dd> *xo, std::vector<dd> *yo, std::vector<dd> *zo) {
    s = ctx.stream;
    m = hp.m;
    nB = (int)hp.bs.size();
    nn = hp.nn;
    nlp = hp.nlp;
    nt = nn + nlp;
    nDim = 0;
    for (int n : hp.bs) nDim += n;
    nDim += (int)nlp;
    dd **bufs[] = {&X, &Z, &invCholX, &invCholZ, &invZ, &DxMat, &DzMat, &r_zinv, &x_rd_zinv, &DLS1, &DLS2, &dualMat};
    for (dd **b : bufs)
        if (!(*b = alloc(nt))) return -2;
    if (!(y = alloc(m)) || !(dy = alloc(m)) || !(g = alloc(m)) || !(primalVec = alloc(m)) ||
        !(bMat = alloc((idx_t)m * m)))
        return -2;
    slot_xz = 3 * nB;
    slot_init = 4 * nB;
    slot_C = 6 * nB;
    slot_b = slot_C + 1;
    slot_C2 = slot_C + 2;
    slot_b2 = slot_C + 3;
    slot_max = slot_C + 4;
    slot_max2 = slot_C + 5;
    slot_bmax = slot_C + 6;
    slot_lp = slot_C + 7;
    nslots = slot_lp + 6 * (int)nlp;
    for (hipEvent_t *e : {&ev_t0, &ev_t1, &ev_t2})
        if (hipEventCreate(e) != hipSuccess) return -2;
    if (!(d_sc = alloc(nslots)) || !(d_diag = alloc(2 * (idx_t)nDim))) return -2;
    if (hipHostMalloc((void **)&h_sc, sizeof(dd) * nslots) != hipSuccess) return -2;
    if (hipHostMalloc((void **)&h_diag, sizeof(dd) * 2 * (size_t)nDim) != hipSuccess) return -2;
    if (hipHostMalloc((void **)&h_infos, sizeof(int64_t) * (2 * nB + 2)) != hipSuccess) return -2;
    if (hipMalloc((void **)&d_infos, sizeof(int64_t) * (2 * nB + 2)) != hipSuccess) return -2;
    mem.push_back(d_infos);
    // Rsyev step-length tasks: (L_X^{-1}, dX) of each small block, then (L_Z^{-1}, dZ)
    for (int l = 0; l < nB; l++)
        if (bsz(l) <= sdk::SMALL_EIG_MAXN) small_blk.push_back(l);
    if (!small_blk.empty()) {
        const int ns = (int)small_blk.size();
        std::vector<sdk::SmallEigTask> tk;
        for (int pass = 0; pass < 2; pass++)
            for (int l : small_blk)
                tk.push_back({(pass ? invCholZ : invCholX) + dp.boff[l], (pass ? DzMat : DxMat) + dp.boff[l], bsz(l)});
        if (hipMalloc((void **)&d_small_tasks, sizeof(sdk::SmallEigTask) * tk.size()) != hipSuccess) return -2;
        mem.push_back(d_small_tasks);
        if (hipMemcpy(d_small_tasks, tk.data(), sizeof(sdk::SmallEigTask) * tk.size(), hipMemcpyHostToDevice) !=
            hipSuccess)
            return -2;
        if (!(d_small_out = alloc(2 * ns))) return -2;
        if (hipMalloc((void **)&d_small_info, sizeof(int) * 2 * ns) != hipSuccess) return -2;
        mem.push_back(d_small_info);
        if (hipHostMalloc((void **)&h_small_out, sizeof(dd) * 2 * ns) != hipSuccess) return -2;
        if (hipHostMalloc((void **)&h_small_info, sizeof(int) * 2 * ns) != hipSuccess) return -2;
    }
    if (nlp) {
        if (!(d_lp_eig = alloc(2 * nlp))) return -2;
        if (hipHostMalloc((void **)&h_lp_eig, sizeof(dd) * 2 * nlp) != hipSuccess) return -2;
    }
    if (!(d_rinv2 = alloc(sdk::POTRF_SCRATCH_DD)) || !(d_rinv3 = alloc(sdk::POTRF_SCRATCH_DD))) return -2;
    if (!spd && ctx.path >= 1 && (!(bSave = alloc((idx_t)m * m)) || !(rvec = alloc(m)) || !(dyTry = alloc(m))))
        return -2;
    if (spd && (!(spv = alloc(spd->nnz)) || !(spx = alloc(m)))) return -2;
    if (hipStreamCreateWithFlags(&s2, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s3, hipStreamNonBlocking) != hipSuccess)
        return -2;
    if (hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ev_join3, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&la.ready, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&la.done, hipEventDisableTiming) != hipSuccess)
        return -2;
    la.hs = s2;
    la.w2 = &ozw2;
    // the panels' planes alternate between ozw2 and ozw3 (s3's X
    // factorisation is joined before the Schur factorisation starts and
    // forked after it ends)
    la.wp = &ozw3;
    {
        std::vector<sdk::DotTask> txz, t3;
        const dd *a3[3] = {X, DxMat, DxMat}, *b3[3] = {DzMat, Z, DzMat};
        for (int l = 0; l < nB; l++) {
            idx_t ne = (idx_t)bsz(l) * bsz(l), o = dp.boff[l];
            max_blk_elems = std::max(max_blk_elems, ne);
            txz.push_back({X + o, Z + o, ne, d_sc + slot_xz + l});
        }
        for (int q = 0; q < 3; q++)
            for (int l = 0; l < nB; l++) {
                idx_t ne = (idx_t)bsz(l) * bsz(l), o = dp.boff[l];
                t3.push_back({a3[q] + o, b3[q] + o, ne, d_sc + q * nB + l});
            }
        if (ctx.path == 1 && !(gsplit = alloc(sdk::GEMM_KS * max_blk_elems))) return -2;
        if (hipMalloc((void **)&d_tasks_xz, sizeof(sdk::DotTask) * txz.size()) != hipSuccess) return -2;
        mem.push_back(d_tasks_xz);
        if (init) {
            if (!(initX = alloc(nt)) || !(initZ = alloc(nt))) return -2;
            std::vector<sdk::DotTask> ti;
            for (int q = 0; q < 2; q++)
                for (int l = 0; l < nB; l++) {
                    idx_t ne = (idx_t)bsz(l) * bsz(l), o = dp.boff[l];
                    ti.push_back({(q ? X : initX) + o, (q ? initZ : Z) + o, ne, d_sc + slot_init + q * nB + l});
                }
            if (!ti.empty()) {
                if (hipMalloc((void **)&d_tasks_init, sizeof(sdk::DotTask) * ti.size()) != hipSuccess) return -2;
                mem.push_back(d_tasks_init);
                if (hipMemcpy(d_tasks_init, ti.data(), sizeof(sdk::DotTask) * ti.size(), hipMemcpyHostToDevice) !=
                    hipSuccess)
                    return -2;
            }
        }
        if (hipMalloc((void **)&d_tasks3, sizeof(sdk::DotTask) * t3.size()) != hipSuccess) return -2;
        mem.push_back(d_tasks3);
        if (hipMemcpy(d_tasks_xz, txz.data(), sizeof(sdk::DotTask) * txz.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_tasks3, t3.data(), sizeof(sdk::DotTask) * t3.size(), hipMemcpyHostToDevice) != hipSuccess)
            return -2;
        if (!(d_partial = alloc((idx_t)std::max(1, 3 * nB) * ctx.dot_T))) return -2;
    }
    double t_start = now();
    for (dd **b : bufs) sdk::fill(s, nt, mk(0.0), *b);

    // Solutions::initialize (sdpa_dataset.cpp:35-60)
    ddr lambda = prm.lambdaStar;
    for (int l = 0; l < nB; l++) {
        int n = bsz(l);
        sdk::set_identity(s, n, lambda.v, blk(X, l));
        sdk::set_identity(s, n, lambda.v, blk(Z, l));
        sdk::set_identity(s, n, (1.0 / sqrt(lambda)).v, blk(invCholX, l));
        sdk::set_identity(s, n, (1.0 / sqrt(lambda)).v, blk(invCholZ, l));
        sdk::set_identity(s, n, (1.0 / lambda).v, blk(invZ, l));
    }
    if (nlp) {  // DenseLinearSpace::setIdentity sets every LP entry to the scalar
        sdk::fill(s, nlp, lambda.v, X + nn);
        sdk::fill(s, nlp, lambda.v, Z + nn);
        sdk::fill(s, nlp, (1.0 / sqrt(lambda)).v, invCholX + nn);
        sdk::fill(s, nlp, (1.0 / sqrt(lambda)).v, invCholZ + nn);
        sdk::fill(s, nlp, (1.0 / lambda).v, invZ + nn);
    }
    sdk::fill(s, m, mk(0.0), y);
    if (init) {
        // sdpa_main.cpp:237-251: the initial point over lambda I, its inverse
        // factors and copies.  (The reference goes on with whatever a failed
        // factorisation leaves; this stops with a message.)
        std::vector<dd> hx, hy, hz;
        init->build(hp, prm.lambdaStar, hx, hy, hz);
        if ((idx_t)hx.size() != nt || (idx_t)hz.size() != nt || (int)hy.size() != m) {
            ctx.set_error("initial point does not match the problem");
            return -1;
        }
        if (hipMemcpyAsync(X, hx.data(), sizeof(dd) * nt, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(Z, hz.data(), sizeof(dd) * nt, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(y, hy.data(), sizeof(dd) * m, hipMemcpyHostToDevice, s) != hipSuccess)
            return -2;
        sync();
        if (!inv_chol_xz()) {
            ctx.set_error("initial point: X or Z is not positive definite");
            return -1;
        }
        sdk::copy(s, nt, X, initX);
        sdk::copy(s, nt, Z, initZ);
    }

    ddr initNormP, initNormD, curNormP, curNormD;
    residuals(initNormP, initNormD);
    curNormP = initNormP;
    curNormD = initNormD;

    ddr alphaP = 0.0, alphaD = 0.0, betaV = prm.betaStar;
    bool switchOn = true;
    ddr muInitial = ddr(prm.lambdaStar) * ddr(prm.lambdaStar), muCurrent = muInitial;
    if (init) {  // AverageComplementarity::initialize(Solutions&) (sdpa_parts.cpp:583-588)
        muInitial = inner_xz();
        muInitial /= nDim;
        muCurrent = muInitial;
    }
    ddr thetaP, thetaD;
    {
        ddr accuracy = prm.epsilonDash;
        thetaP = (initNormP < accuracy) ? ddr(0.0) : ddr(1.0);
        thetaD = (initNormD < accuracy) ? ddr(0.0) : ddr(1.0);
    }
    ddr rho = 1.0, etaPrimal, etaDual, objValPrimal, objValDual;
    {
        ddr omegaStar = prm.omegaStar;
        etaPrimal = omegaStar * nDim * muInitial;
        etaDual = omegaStar * nDim * muInitial;
        objValPrimal = inner_C(X);
        objValDual = inner_b(y);
    }
    int phase = noINFO;
    auto updateCheck = [&](const ddr &nP, const ddr &nD) -> bool {
        // Phase::updateCheck (sdpa_parts.cpp:875-954)
        const ddr NONZERO = 1.0e-6;
        ddr accuracy = prm.epsilonDash;
        phase = noINFO;
        if (nP <= accuracy) phase = (nD <= accuracy) ? pdFEAS : pFEAS;
        if (phase == noINFO && nD <= accuracy) phase = dFEAS;
        if (phase == pdFEAS) {
            ddr mean = (abs(objValPrimal) + abs(objValDual)) / 2.0;
            ddr PDgap = abs(objValPrimal - objValDual);
            ddr dominator = (mean < 1.0) ? ddr(1.0) : mean;
            if (PDgap / dominator <= prm.epsilonStar) {
                phase = pdOPT;
                return false;
            }
        }
        if (phase == noINFO && rho > 1.0 + NONZERO) {
            phase = pdINF;
            return false;
        }
        if (phase == pFEAS) {
            if (objValPrimal <= -prm.upperBound) {
                phase = pUNBD;
                return false;
            }
            if (rho > 1.0 + NONZERO) {
                phase = pFEAS_dINF;
                return false;
            }
        }
        if (phase == dFEAS) {
            if (objValDual >= -prm.lowerBound) {
                phase = dUNBD;
                return false;
            }
            if (rho > 1.0 + NONZERO) {
                phase = pINF_dFEAS;
                return false;
            }
        }
        return true;
    };
    updateCheck(initNormP, initNormD);

    // Newton::Mehrotra (sdpa_newton.cpp:1345-1384)
    auto mehrotra = [&](bool predictor) -> bool {
        ddr target = betaV * muCurrent;
        sdk::copy(s, nt, invZ, r_zinv);
        sdk::scal(s, nt, target.v, r_zinv);
        plus_in(r_zinv, X, mk(-1.0));
        if (!predictor) {
            triple(DLS1, DxMat, DzMat, invZ, DLS2);
            plus_in(r_zinv, DLS1, mk(-1.0));
        }
        if (phase == pFEAS || phase == noINFO) {
            if (predictor) triple(x_rd_zinv, X, dualMat, invZ, DLS2);
            plus(DLS1, r_zinv, x_rd_zinv, mk(-1.0));
        } else {
            sdk::copy(s, nt, r_zinv, DLS1);
        }
        sdp::inner_A(s, dp, DLS1, g, ctx.dot_T, ctx.path >= 1);
        sdk::scal(s, m, mk(-1.0), g);
        sdk::axpy(s, m, mk(1.0), primalVec, g);
        if (predictor) {
            // Schur build and Cholesky timed with events; one synchronisation
            // brings back the Cholesky's info (and max|B| for the refinement)
            (void)hipEventRecord(ev_t0, s);
            // the sparse branch assembles in the reference's sparse LP order
            sdp::schur(s, dp, X, invZ, bMat, ctx.dot_T, spd != nullptr);
            if (spd) sps::gather(s, *spd, bMat, spv);
            (void)hipEventRecord(ev_t1, s);
            int64_t info = 0;
            if (spd) {
                // Lal::getCholesky on the pattern: no failure exit
                sps::factor(s, *spd, spv);
                (void)hipEventRecord(ev_t2, s);
            } else {
                if (ctx.path >= 1) {
                    // bSave = B and max|B| in one pass over B
                    sdk::max_abs(s, (idx_t)m * m, bMat, d_sc + slot_bmax, ctx.d_max_work, ctx.d_counter + 1, bSave);
                    (void)hipMemcpyAsync(h_sc + slot_bmax, d_sc + slot_bmax, sizeof(dd), hipMemcpyDeviceToHost, s);
                }
                potrf(s, m, bMat, ctx.d_info, ctx.d_rinv, ctx.ozw);
                (void)hipEventRecord(ev_t2, s);
                (void)hipMemcpyAsync(&info, ctx.d_info, sizeof(info), hipMemcpyDeviceToHost, s);
            }
            // the Ozaki form solves with the block-inverse operators of this
            // factor: enqueued before the host reads info, so the device goes
            // on (after a failed factorisation they are never used)
            const bool want_binv = !spd && ctx.path >= 1 && m >= 512;
            const bool binv_ok = want_binv && ctx.binv_prepare(s, m, bMat, m) == 0;
            sync();
            float ms1 = 0, ms2 = 0;
            if (hipEventElapsedTime(&ms1, ev_t0, ev_t1) == hipSuccess) t_schur += 1e-3 * ms1;
            if (hipEventElapsedTime(&ms2, ev_t1, ev_t2) == hipSuccess) t_chol += 1e-3 * ms2;
            if (ctx.path >= 1 && !spd) bmax = sc(slot_bmax);
            if (info == sdk::POTRF_WAIT_FAILED) wait_failed = true;
            if (info != 0) return false;
            binv = binv_ok;
        }
        schur_solve();
        sdp::plus_A(s, dp, dy, DzMat);
        sdk::scal(s, nt, mk(-1.0), DzMat);
        if (phase == pFEAS || phase == noINFO) plus_in(DzMat, dualMat, mk(1.0));
        triple(DLS1, X, DzMat, invZ, DLS2);
        plus(DxMat, r_zinv, DLS1, mk(-1.0));
        for (int l = 0; l < nB; l++) sdk::symmetrize(s, bsz(l), blk(DxMat, l));
        return true;
    };

    int pIteration = 0;
    while (updateCheck(curNormP, curNormD) && pIteration < prm.maxIteration) {
        switchOn = (phase == noINFO || phase == pFEAS || phase == dFEAS);
        if (phase == pdFEAS) {
            betaV = 0.0;
        } else {
            betaV = prm.betaBar;
            if (!switchOn) betaV = ddr(2.0);
        }
        if (!mehrotra(true) || oz_failed) {
            res.message = oz_failed ? "Ozaki workspace allocation failed" : "Schur Cholesky failed";
            break;
        }
        // StepLength::MehrotraPredictor (sdpa_parts.cpp:307-354) and
        // DirectionParameter::MehrotraCorrector read C . dX, b . dy and the
        // three direction products: one batch, one synchronisation
        enq_inner_C(DxMat, slot_C);
        enq_inner_b(dy, slot_b);
        enq_dirs();
        fetch_all();
        alphaP = alphaD = 0.9;
        if (phase == noINFO || phase == dFEAS) {
            if (alphaP > 1.0) alphaP = 1.0;
        } else {
            ddr inc = sc(slot_C);
            if (inc > 0.0) {
                if (alphaP > alphaD) alphaP = alphaD;
                if (alphaP > 1.0) alphaP = 1.0;
            }
        }
        if (phase == noINFO || phase == pFEAS) {
            if (alphaD > 1.0) alphaD = 1.0;
        } else {
            ddr inc = sc(slot_b);
            if (inc < 0.0) {
                if (alphaD > alphaP) alphaD = alphaP;
                if (alphaD > 1.0) alphaD = 1.0;
            }
        }
        // DirectionParameter::MehrotraCorrector (sdpa_parts.cpp:499-529)
        {
            ddr xv, uz, uv;
            read_dirs(xv, uz, uv);
            ddr muTarget = muCurrent + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nDim;
            betaV = muTarget / muCurrent;
            if (betaV < 1.0) betaV = betaV * betaV;
            if (phase == pdFEAS) {
                if (betaV < prm.betaStar) betaV = prm.betaStar;
                if (betaV > 1.0) betaV = 1.0;
            } else {
                if (betaV < prm.betaBar) betaV = prm.betaBar;
            }
        }
        if ((phase == pdFEAS) && ((betaV > 5) || (objValPrimal < objValDual))) break;
        mehrotra(false);
        // StepLength::MehrotraCorrector (sdpa_parts.cpp:356-466)
        {
            const ddr xi = 3.0;
            ddr alphaBD = 100.0;
            ddr mx, mz;
            // every inner product this step may read, downloaded ahead of the
            // eigenvalue solves' synchronisation
            enq_inner_C(DxMat, slot_C);
            enq_inner_b(dy, slot_b);
            enq_dirs();
            enq_inner_b(y, slot_b2);
            enq_inner_C(X, slot_C2);
            prefetch_all();
            bool eig_ok = min_eigen_xz(mx, mz);  // synchronises: h_sc is current
            if (!eig_ok) {
                res.message = "eigenvalue iteration did not converge";
                break;
            }
            alphaP = (-mx > 1.0 / alphaBD) ? -1.0 / mx : alphaBD;
            alphaD = (-mz > 1.0 / alphaBD) ? -1.0 / mz : alphaBD;
            alphaP = prm.gammaStar * alphaP;
            alphaD = prm.gammaStar * alphaD;
            if (phase == noINFO || phase == dFEAS) {
                if (alphaP > 1.0) alphaP = 1.0;
            } else {
                ddr inc = sc(slot_C);
                if (inc > 0.0) {
                    if (alphaP > alphaD) alphaP = alphaD;
                    if (alphaP > 1.0) alphaP = 1.0;
                }
            }
            if (phase == noINFO || phase == pFEAS) {
                if (alphaD > 1.0) alphaD = 1.0;
            } else {
                ddr inc = sc(slot_b);
                if (inc < 0.0) {
                    if (alphaD > alphaP) alphaD = alphaP;
                    if (alphaD > 1.0) alphaD = 1.0;
                }
            }
            if (switchOn && (phase == noINFO || phase == pFEAS || phase == dFEAS)) {
                ddr xv, uz, uv;
                read_dirs(xv, uz, uv);
                ddr thetaMax = max((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
                ddr muNew = muCurrent + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nDim;
                while (thetaMax * thetaMax * muInitial > xi * muNew) {
                    ddr alphaMax = 0.95 * max(alphaP, alphaD);
                    alphaP = min(alphaP, alphaMax);
                    alphaD = min(alphaD, alphaMax);
                    thetaMax = max((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
                    muNew = muCurrent + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nDim;
                    if (alphaP < 1.0e-6 && alphaD < 1.0e-6) break;
                }
            }
            if (phase == pdFEAS) {
                ddr oD = sc(slot_b2);
                ddr oP = sc(slot_C2);
                ddr incD = sc(slot_b);
                incD *= alphaD;
                ddr incP = sc(slot_C);
                incP *= alphaP;
                ddr maxRatio = (oD - oP) / (incP - incD);
                if ((maxRatio > 0.0) && (maxRatio < 1.0)) {
                    alphaP *= maxRatio;
                    alphaD *= maxRatio;
                }
            }
        }
        res.trace.push_back({pIteration, muCurrent.v, thetaP.v, thetaD.v, objValPrimal.v, objValDual.v, alphaP.v, alphaD.v,
                             betaV.v});
        // Solutions::update (sdpa_dataset.cpp:120-143)
        plus_in(X, DxMat, alphaP.v);
        sdk::axpy(s, m, alphaD.v, dy, y);
        plus_in(Z, DzMat, alphaD.v);
        const ddr cannot_move = 1.0e-4;
        if (alphaP < cannot_move && alphaD < cannot_move) {
            res.message = "cannot move";
            break;
        }
        // the new point's inverse factors, mu, residuals, objectives and
        // SolveInfo products: one batch, one synchronisation (none of the
        // values is used when a factorisation failed)
        {
            double t0 = now();
            enq_inv_chol_fork();
            enq_xz();
            enq_residuals();
            enq_inner_C(X, slot_C2);
            enq_inner_b(y, slot_b2);
            if (init)
                enq_init();
            else
                enq_diag();
            enq_inv_chol_launch();
            enq_inv_chol_join();
            fetch_all();
            t_inverse += now() - t0;
        }
        if (!inv_chol_ok()) {
            res.message = "cannot move";
            break;
        }
        if (switchOn) {
            thetaP = abs((1.0 - alphaP) * thetaP);
            thetaD = abs((1.0 - alphaD) * thetaD);
        }
        muCurrent = read_xz();
        muCurrent /= nDim;
        curNormP = sc(slot_max);
        curNormD = sc(slot_max2);
        thetaP = (initNormP > 1.0e-10) ? curNormP / initNormP : ddr(0.0);
        thetaD = (initNormD > 1.0e-10) ? curNormD / initNormD : ddr(0.0);
        {
            // SolveInfo::update (sdpa_parts.cpp:743-790; with an initial point 691-740)
            objValPrimal = sc(slot_C2);
            objValDual = sc(slot_b2);
            ddr primal = thetaP, dual = thetaD, omega = prm.omegaStar;
            rho = 0.0;
            ddr x0z0 = nDim * muInitial, xMatzMat = nDim * muCurrent, x0zMat = 0.0, xMatz0 = 0.0;
            if (init)
                read_init(x0zMat, xMatz0);
            else
                read_diag(prm.lambdaStar, x0zMat, xMatz0);
            ddr accuracy = prm.epsilonDash;
            if (curNormP <= accuracy && xMatz0 < etaPrimal) etaPrimal = xMatz0;
            if (curNormD <= accuracy && x0zMat < etaDual) etaDual = x0zMat;
            if (curNormP > accuracy && curNormD <= accuracy)
                rho = primal * x0zMat / ((primal + (1.0 - primal) * omega) * etaDual + xMatzMat);
            if (curNormP <= accuracy && curNormD > accuracy)
                rho = dual * xMatz0 / ((dual + (1.0 - dual) * omega) * etaPrimal + xMatzMat);
            if (curNormP > accuracy && curNormD > accuracy)
                rho = (dual * xMatz0 + primal * x0zMat) /
                      ((primal * dual + omega * (primal * (1.0 - dual) + (1.0 - primal) * dual)) * x0z0 + xMatzMat);
        }
        pIteration++;
    }
    residuals(curNormP, curNormD);
    sync();
    res.t.total = now() - t_start;
    res.t.schur = t_schur;
    res.t.chol = t_chol;
    res.t.lanczos = t_lanczos;
    res.t.inverse = t_inverse;
    res.refine_sweeps = refine_accepted;
    res.iterations = pIteration;
    res.phase = phase;
    int rep = phase;
    switch (phase) {  // Phase::reverse
    case pFEAS: rep = dFEAS; break;
    case dFEAS: rep = pFEAS; break;
    case pFEAS_dINF: rep = pINF_dFEAS; break;
    case pINF_dFEAS: rep = pFEAS_dINF; break;
    case pUNBD: rep = dUNBD; break;
    case dUNBD: rep = pUNBD; break;
    default: break;
    }
    res.phase_reported = rep;
    res.mu = muCurrent.v;
    res.objValPrimal = objValPrimal.v;
    res.objValDual = objValDual.v;
    res.normPrimalVec = curNormP.v;
    res.normDualMat = curNormD.v;
    {
        ddr mean = (abs(objValPrimal) + abs(objValDual)) / 2.0;
        ddr PDgap = abs(objValPrimal - objValDual);
        res.relgap = ((mean < 1.0) ? PDgap : PDgap / mean).v;
    }
    if (xo) {
        xo->resize(nt);
        (void)hipMemcpy(xo->data(), X, sizeof(dd) * nt, hipMemcpyDeviceToHost);
    }
    if (zo) {
        zo->resize(nt);
        (void)hipMemcpy(zo->data(), Z, sizeof(dd) * nt, hipMemcpyDeviceToHost);
    }
    if (yo) {
        yo->resize(m);
        (void)hipMemcpy(yo->data(), y, sizeof(dd) * m, hipMemcpyDeviceToHost);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx.set_error(hipGetErrorString(e));
        return -3;
    }
    if (wait_failed) {
        ctx.set_error("solve: a Cholesky panel wait for its diagonal block ran out of its bound");
        return -3;
    }
    return 0;
}

}  // namespace

int solve(Context &ctx, const HostProblem &hp, DeviceProblem &dp, const Params &prm, SolveOutput &out, std::vector<dd> *xMat,
          std::vector<dd> *yVec, std::vector<dd> *zMat, const InitPoint *init, const sps::Device *sparse) {
    Solver sv(ctx, hp, dp, prm, init, sparse && sparse->on ? sparse : nullptr);
    return sv.run(out, xMat, yVec, zMat);
}

dd lanczos_min_eigen(Context &ctx, int n, const dd *d_L, const dd *d_M) {
    ChainPool pool;
    if (!pool.ensure(1, n)) return dd{NAN, NAN};
    pool.chains[0].n = n;
    pool.chains[0].L = d_L;
    pool.chains[0].M = d_M;
    run_chains(pool, 1, ctx.dot_T, ctx.stream, false);
    (void)hipStreamSynchronize(ctx.stream);
    return pool.chains[0].result.v;
}

}  // namespace sdp
