// The sparse-Schur path: host analysis (decision, ordering, pattern,
// schedules) and the single-workgroup kernels that factor and solve on the
// pattern in the reference's per-entry operation order.
#include "sparse_schur.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace sps {

// ------------------------------------------------------------ analysis --
std::vector<int> min_degree(int m, const std::vector<std::vector<int>> &adj, std::vector<std::vector<int>> &rows) {
    const int W = (m + 63) / 64;
    std::vector<uint64_t> g((size_t)m * W, 0), alive(W, 0);
    auto rowp = [&](int i) { return g.data() + (size_t)i * W; };
    for (int i = 0; i < m; i++) {
        alive[i >> 6] |= uint64_t(1) << (i & 63);
        for (int j : adj[i])
            if (j != i) {
                rowp(i)[j >> 6] |= uint64_t(1) << (j & 63);
                rowp(j)[i >> 6] |= uint64_t(1) << (i & 63);
            }
    }
    std::vector<int> deg(m);
    for (int i = 0; i < m; i++) {
        int c = 0;
        for (int w = 0; w < W; w++) c += __builtin_popcountll(rowp(i)[w]);
        deg[i] = c;
    }
    std::vector<int> order, newIdx(m, -1);
    std::vector<std::vector<int>> nb_old(m);
    std::vector<uint64_t> nb(W);
    for (int step = 0; step < m; step++) {
        int v = -1;
        for (int i = 0; i < m; i++)
            if (((alive[i >> 6] >> (i & 63)) & 1) && (v < 0 || deg[i] < deg[v])) v = i;
        order.push_back(v);
        newIdx[v] = step;
        alive[v >> 6] &= ~(uint64_t(1) << (v & 63));
        for (int w = 0; w < W; w++) nb[w] = rowp(v)[w] & alive[w];
        for (int w = 0; w < W; w++)
            for (uint64_t bits = nb[w]; bits; bits &= bits - 1) {
                const int u = w * 64 + __builtin_ctzll(bits);
                nb_old[step].push_back(u);
                uint64_t *ru = rowp(u);
                int c = 0;
                for (int q = 0; q < W; q++) {
                    ru[q] = (ru[q] | nb[q]) & alive[q];
                    c += __builtin_popcountll(ru[q]);
                }
                ru[u >> 6] &= ~(uint64_t(1) << (u & 63));
                deg[u] = c - 1;  // its own bit came in with nb
            }
    }
    rows.assign(m, {});
    for (int p = 0; p < m; p++) {
        for (int u : nb_old[p]) rows[p].push_back(newIdx[u]);
        std::sort(rows[p].begin(), rows[p].end());
    }
    return order;
}

void analyse(const sdp::HostProblem &P, int mode, Analysis &S) {
    S = Analysis();
    if (mode == 1) return;
    const int m = P.m;
    const int nBlock2 = (int)P.bs.size() + P.nlp;
    const int m_thr = mode == 2 ? 0 : 100, b_thr = mode == 2 ? 0 : 5;
    const double agg = mode == 2 ? 2.0 : 0.25, ext = mode == 2 ? 2.0 : 0.4;
    if (m <= m_thr || nBlock2 <= b_thr) return;
    for (const auto &c : P.cons)
        if ((int)c.size() > m * std::sqrt(agg)) return;
    for (const auto &c : P.lp_cons)
        if ((int)c.size() > m * std::sqrt(agg)) return;
    // Chordal::makeGraph: constraints sharing an SDP block or an LP entry
    std::vector<std::vector<int>> adj(m);
    auto add = [&](const std::vector<int> &cs) {
        for (int i : cs) adj[i].insert(adj[i].end(), cs.begin(), cs.end());
    };
    for (const auto &c : P.cons) add(c);
    for (const auto &c : P.lp_cons) add(c);
    long long tsize = 0;
    for (auto &a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        tsize += (long long)a.size();
    }
    S.aggregate = tsize;
    if (tsize > agg * m * m) return;
    std::vector<std::vector<int>> rows;
    std::vector<int> order = min_degree(m, adj, rows);
    long long nonzeros = 0;
    for (const auto &r : rows) nonzeros += 1 + (long long)r.size();
    if (nonzeros * 2 - m > ext * m * m) return;

    S.on = true;
    S.m = m;
    S.ordering = order;
    S.reverse.assign(m, 0);
    for (int i = 0; i < m; i++) S.reverse[order[i]] = i;
    S.diag.assign(m + 1, 0);
    for (int p = 0; p < m; p++) S.diag[p + 1] = S.diag[p] + 1 + (int)rows[p].size();
    const int nnz = S.diag[m];
    S.col.resize(nnz);
    S.row.resize(nnz);
    for (int p = 0; p < m; p++) {
        int k = S.diag[p];
        S.col[k] = p;
        S.row[k] = p;
        for (int c : rows[p]) {
            ++k;
            S.col[k] = c;
            S.row[k] = p;
        }
    }
    S.gather.resize(nnz);
    for (int k = 0; k < nnz; k++) S.gather[k] = (idx_t)S.ordering[S.row[k]] + (idx_t)m * S.ordering[S.col[k]];

    // the row loop of getCholesky, run symbolically: row i updates entry
    // (c1, c2) of row c1 with U_{i,c1} U_{i,c2} for every pair k1 <= k2 of
    // its off-diagonal entries; the lists come out with i ascending
    std::vector<int> cnt(nnz + 1, 0);
    auto walk = [&](bool fill) {
        for (int i = 0; i < m; i++)
            for (int k1 = S.diag[i] + 1; k1 < S.diag[i + 1]; k1++) {
                const int c1 = S.col[k1];
                int k3 = S.diag[c1];
                for (int k2 = k1; k2 < S.diag[i + 1]; k2++) {
                    const int c2 = S.col[k2];
                    while (k3 < S.diag[c1 + 1] && S.col[k3] != c2) k3++;
                    if (k3 == S.diag[c1 + 1]) break;  // cannot happen: the pattern is closed under elimination
                    if (fill) {
                        const int q = cnt[k3]++;
                        S.con_a[q] = k1;
                        S.con_b[q] = k2;
                    } else {
                        cnt[k3 + 1]++;
                    }
                    k3++;
                }
            }
    };
    walk(false);
    for (int k = 0; k < nnz; k++) cnt[k + 1] += cnt[k];
    S.con_ptr.assign(cnt.begin(), cnt.end());
    S.con_a.resize(cnt[nnz]);
    S.con_b.resize(cnt[nnz]);
    walk(true);  // cnt[k] runs from con_ptr[k] up to con_ptr[k + 1]

    // levels: a row waits for every row whose pattern holds its index
    std::vector<int> lev(m, 0);
    int nlev = 0;
    for (int i = 0; i < m; i++) {
        for (int k = S.diag[i] + 1; k < S.diag[i + 1]; k++) lev[S.col[k]] = std::max(lev[S.col[k]], lev[i] + 1);
        nlev = std::max(nlev, lev[i] + 1);
    }
    S.lev_ptr.assign(nlev + 1, 0);
    for (int i = 0; i < m; i++) S.lev_ptr[lev[i] + 1]++;
    for (int L = 0; L < nlev; L++) S.lev_ptr[L + 1] += S.lev_ptr[L];
    S.lev_rows.resize(m);
    {
        std::vector<int> pos(S.lev_ptr.begin(), S.lev_ptr.end() - 1);
        for (int i = 0; i < m; i++) S.lev_rows[pos[lev[i]]++] = i;
    }
    S.lev_ent_ptr.assign(nlev + 1, 0);
    for (int L = 0; L < nlev; L++) {
        for (int r = S.lev_ptr[L]; r < S.lev_ptr[L + 1]; r++) {
            const int i = S.lev_rows[r];
            for (int k = S.diag[i]; k < S.diag[i + 1]; k++) S.lev_ents.push_back(k);
        }
        S.lev_ent_ptr[L + 1] = (int)S.lev_ents.size();
    }
    // forward solve: the rows that update x_j, ascending
    S.fw_ptr.assign(m + 1, 0);
    for (int k = 0; k < nnz; k++)
        if (S.col[k] != S.row[k]) S.fw_ptr[S.col[k] + 1]++;
    for (int j = 0; j < m; j++) S.fw_ptr[j + 1] += S.fw_ptr[j];
    S.fw_row.resize(S.fw_ptr[m]);
    S.fw_pos.resize(S.fw_ptr[m]);
    {
        std::vector<int> pos(S.fw_ptr.begin(), S.fw_ptr.end() - 1);
        for (int k = 0; k < nnz; k++)
            if (S.col[k] != S.row[k]) {
                const int q = pos[S.col[k]]++;
                S.fw_row[q] = S.row[k];
                S.fw_pos[q] = k;
            }
    }
    // backward levels: a row waits for the rows of its pattern
    std::vector<int> blev(m, 0);
    int nb = 0;
    for (int i = m - 1; i >= 0; i--) {
        for (int k = S.diag[i] + 1; k < S.diag[i + 1]; k++) blev[i] = std::max(blev[i], blev[S.col[k]] + 1);
        nb = std::max(nb, blev[i] + 1);
    }
    S.bw_ptr.assign(nb + 1, 0);
    for (int i = 0; i < m; i++) S.bw_ptr[blev[i] + 1]++;
    for (int L = 0; L < nb; L++) S.bw_ptr[L + 1] += S.bw_ptr[L];
    S.bw_rows.resize(m);
    {
        std::vector<int> pos(S.bw_ptr.begin(), S.bw_ptr.end() - 1);
        for (int i = 0; i < m; i++) S.bw_rows[pos[blev[i]]++] = i;
    }
}

// -------------------------------------------------------------- device --
template <class T>
static T *dcopy(Device &D, const std::vector<T> &v, hipStream_t s, bool &ok) {
    T *p = nullptr;
    const size_t bytes = sizeof(T) * std::max<size_t>(1, v.size());
    if (hipMalloc((void **)&p, bytes) != hipSuccess) {
        ok = false;
        return nullptr;
    }
    D.allocs.push_back(p);
    if (!v.empty()) (void)hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s);
    return p;
}

int upload(const Analysis &S, Device &D, hipStream_t s, std::string &err) {
    release(D);
    if (!S.on) return 0;
    bool ok = true;
    D.m = S.m;
    D.nnz = S.diag[S.m];
    D.nlev = (int)S.lev_ptr.size() - 1;
    D.nblev = (int)S.bw_ptr.size() - 1;
    D.d_ordering = dcopy(D, S.ordering, s, ok);
    D.d_diag = dcopy(D, S.diag, s, ok);
    D.d_col = dcopy(D, S.col, s, ok);
    D.d_row = dcopy(D, S.row, s, ok);
    D.d_gather = dcopy(D, S.gather, s, ok);
    D.d_lev_ptr = dcopy(D, S.lev_ptr, s, ok);
    D.d_lev_rows = dcopy(D, S.lev_rows, s, ok);
    D.d_lev_ent_ptr = dcopy(D, S.lev_ent_ptr, s, ok);
    D.d_lev_ents = dcopy(D, S.lev_ents, s, ok);
    D.d_con_ptr = dcopy(D, S.con_ptr, s, ok);
    D.d_con_a = dcopy(D, S.con_a, s, ok);
    D.d_con_b = dcopy(D, S.con_b, s, ok);
    D.d_fw_ptr = dcopy(D, S.fw_ptr, s, ok);
    D.d_fw_row = dcopy(D, S.fw_row, s, ok);
    D.d_fw_pos = dcopy(D, S.fw_pos, s, ok);
    D.d_bw_ptr = dcopy(D, S.bw_ptr, s, ok);
    D.d_bw_rows = dcopy(D, S.bw_rows, s, ok);
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
        release(D);
        err = "sparse Schur: device allocation failed";
        return -2;
    }
    D.on = true;
    return 0;
}

void release(Device &D) {
    for (void *p : D.allocs) (void)hipFree(p);
    D = Device();
}

__global__ void k_sps_gather(int nnz, const idx_t *gather, const dd *B, dd *sp) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) sp[k] = B[gather[k]];
}

static constexpr int SPS_NT = 1024;

// Lal::getCholesky per level of rows: every entry of the level takes its
// updates (U_{i,c1} U_{i,c2}, rows i ascending), then each row's diagonal
// becomes 1/sqrt (0 for a negative pivot), then its other entries are
// scaled by it ("*=").  Entries of one level belong to rows that no other
// row of the level updates, so the passes of a level run in parallel.
__global__ __launch_bounds__(SPS_NT) void k_sps_factor(int nlev, const int *lev_ptr, const int *lev_rows,
                                                       const int *lev_ent_ptr, const int *lev_ents, const int *diag,
                                                       const int *row, const int *con_ptr, const int *con_a,
                                                       const int *con_b, dd *sp) {
    const int tid = threadIdx.x;
    for (int L = 0; L < nlev; L++) {
        for (int e = lev_ent_ptr[L] + tid; e < lev_ent_ptr[L + 1]; e += SPS_NT) {
            const int k = lev_ents[e];
            dd v = sp[k];
            for (int q = con_ptr[k]; q < con_ptr[k + 1]; q++) v = dd_sub(v, dd_mul(sp[con_a[q]], sp[con_b[q]]));
            sp[k] = v;
        }
        __syncthreads();
        for (int r = lev_ptr[L] + tid; r < lev_ptr[L + 1]; r += SPS_NT) {
            const int k = diag[lev_rows[r]];
            const dd v = sp[k];
            sp[k] = (v.hi < 0.0 || (v.hi == 0.0 && v.lo < 0.0)) ? mk(0.0) : dd_div(mk(1.0), dd_sqrt(v));
        }
        __syncthreads();
        for (int e = lev_ent_ptr[L] + tid; e < lev_ent_ptr[L + 1]; e += SPS_NT) {
            const int k = lev_ents[e], d = diag[row[k]];
            if (k != d) sp[k] = dd_mul_eq(sp[k], sp[d]);
        }
        __syncthreads();
    }
}

// permuteVec, the forward sweep (x_j: the updates -= U_ij x_i of the rows i
// before it, ascending, then *= its diagonal), the backward sweep (x_i: -=
// U_ij x_j over its pattern, columns descending, then *= its diagonal) and
// reverse_permuteVec; a thread per row of a level.
__global__ __launch_bounds__(SPS_NT) void k_sps_solve(int m, int nlev, const int *lev_ptr, const int *lev_rows,
                                                      int nblev, const int *bw_ptr, const int *bw_rows,
                                                      const int *diag, const int *col, const int *fw_ptr,
                                                      const int *fw_row, const int *fw_pos, const int *ordering,
                                                      const dd *sp, const dd *g, dd *dy, dd *x) {
    const int tid = threadIdx.x;
    for (int k = tid; k < m; k += SPS_NT) x[k] = g[ordering[k]];
    __syncthreads();
    for (int L = 0; L < nlev; L++) {
        for (int r = lev_ptr[L] + tid; r < lev_ptr[L + 1]; r += SPS_NT) {
            const int j = lev_rows[r];
            dd v = x[j];
            for (int q = fw_ptr[j]; q < fw_ptr[j + 1]; q++) v = dd_sub(v, dd_mul(sp[fw_pos[q]], x[fw_row[q]]));
            x[j] = dd_mul_eq(v, sp[diag[j]]);
        }
        __syncthreads();
    }
    for (int L = 0; L < nblev; L++) {
        for (int r = bw_ptr[L] + tid; r < bw_ptr[L + 1]; r += SPS_NT) {
            const int i = bw_rows[r];
            dd v = x[i];
            for (int k = diag[i + 1] - 1; k > diag[i]; k--) v = dd_sub(v, dd_mul(sp[k], x[col[k]]));
            x[i] = dd_mul_eq(v, sp[diag[i]]);
        }
        __syncthreads();
    }
    for (int k = tid; k < m; k += SPS_NT) dy[ordering[k]] = x[k];
}

void gather(hipStream_t s, const Device &D, const dd *B, dd *sp) {
    if (D.nnz > 0) k_sps_gather<<<(D.nnz + 255) / 256, 256, 0, s>>>(D.nnz, D.d_gather, B, sp);
}
void factor(hipStream_t s, const Device &D, dd *sp) {
    sdk::StatScope st("potrf", s, (double)D.nnz);
    k_sps_factor<<<1, SPS_NT, 0, s>>>(D.nlev, D.d_lev_ptr, D.d_lev_rows, D.d_lev_ent_ptr, D.d_lev_ents, D.d_diag,
                                       D.d_row, D.d_con_ptr, D.d_con_a, D.d_con_b, sp);
}
void solve(hipStream_t s, const Device &D, const dd *sp, const dd *g, dd *dy, dd *x) {
    sdk::StatScope st("trsv", s, 2.0 * D.nnz);
    k_sps_solve<<<1, SPS_NT, 0, s>>>(D.m, D.nlev, D.d_lev_ptr, D.d_lev_rows, D.nblev, D.d_bw_ptr, D.d_bw_rows,
                                      D.d_diag, D.d_col, D.d_fw_ptr, D.d_fw_row, D.d_fw_pos, D.d_ordering, sp, g, dy, x);
}

}  // namespace sps
