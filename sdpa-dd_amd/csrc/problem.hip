// Problem input (SDPA sparse format reader, param.sdpa), the Schur formula
// table, the device image of the problem and the kernels that touch the
// constraint matrices: A_i . W inner products, sum_i y_i A_i, and the Schur
// complement assembly with formulas F1/F2/F3.
#include "problem.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <cstring>
#include <fstream>
#include <sstream>

namespace sdp {

using sdk::DotTask;

// ------------------------------------------------------------------ input --
// QD's decimal reader (qd/src/dd_real.cpp read()): r = r*10 + digit in dd,
// then r *= 10^e with binary powering.
static dd npwr10(int n) {
    dd r = mk(10.0), s = mk(1.0);
    int N = n < 0 ? -n : n;
    if (N > 1) {
        while (N > 0) {
            if (N % 2 == 1) s = dd_mul_eq(s, r);
            N /= 2;
            if (N > 0) r = dd_sqr(r);
        }
    } else {
        s = r;
    }
    if (n < 0) return dd_div(mk(1.0), s);
    return s;
}
bool qd_read(const char *s, dd &out) {
    const char *p = s;
    int sign = 0, point = -1, nd = 0, e = 0;
    bool done = false;
    dd r = mk(0.0);
    while (*p == ' ') p++;
    char ch;
    while (!done && (ch = *p) != '\0') {
        if (ch >= '0' && ch <= '9') {
            r = dd_mul_d(r, 10.0);
            r = dd_add_d(r, (double)(ch - '0'));
            nd++;
        } else if (ch == '.') {
            if (point >= 0) return false;
            point = nd;
        } else if (ch == '-' || ch == '+') {
            if (sign != 0 || nd > 0) return false;
            sign = (ch == '-') ? -1 : 1;
        } else if (ch == 'e' || ch == 'E') {
            if (std::sscanf(p + 1, "%d", &e) != 1) return false;
            done = true;
        } else {
            return false;
        }
        p++;
    }
    if (point >= 0) e -= (nd - point);
    if (e != 0) r = dd_mul_eq(r, npwr10(e));
    out = (sign == -1) ? dd_neg(r) : r;
    return true;
}

void Block::change_to_dense() {
    if (dense) return;
    if (nnz_count < (n * n) * 0.20) return;  // sdpa_struct.cpp:465
    dense = true;
    dval.assign((size_t)n * n, mk(0.0));
    for (int k = 0; k < nnz_count; k++) {
        int i = row[k], j = col[k];
        dval[i + (size_t)n * j] = val[k];
        if (i != j) dval[j + (size_t)n * i] = val[k];
    }
    nnz_count = nnz_effect = n * n;
    row.clear();
    col.clear();
    val.clear();
}

namespace {
bool is_numch(char c) { return (c >= '0' && c <= '9') || c == '+' || c == '-'; }
struct Cursor {
    const std::string &s;
    size_t p = 0;
    explicit Cursor(const std::string &x) : s(x) {}
    // fscanf "%*[^0-9+-]": consume one or more characters outside the set
    bool sep() {
        size_t q = p;
        while (q < s.size() && !is_numch(s[q])) q++;
        if (q == p || q >= s.size()) return false;
        p = q;
        return true;
    }
    bool integer(int &v) {
        while (p < s.size() && std::isspace((unsigned char)s[p])) p++;
        size_t q = p;
        if (q < s.size() && (s[q] == '+' || s[q] == '-')) q++;
        size_t d = q;
        while (q < s.size() && std::isdigit((unsigned char)s[q])) q++;
        if (q == d) return false;
        v = std::atoi(s.substr(p, q - p).c_str());
        p = q;
        return true;
    }
    // "%[^,} \t\n]"
    bool token(std::string &t) {
        size_t q = p;
        while (q < s.size() && s[q] != ',' && s[q] != '}' && s[q] != ' ' && s[q] != '\t' && s[q] != '\n') q++;
        if (q == p) return false;
        t = s.substr(p, q - p);
        p = q;
        return true;
    }
};
dd value_of(const std::string &t) {
    dd v;
    if (!qd_read(t.c_str(), v)) v = dd{NAN, NAN};
    return v;
}
}  // namespace

bool read_dat_s(const std::string &path, HostProblem &P, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open data file " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string buf = ss.str();
    Cursor c(buf);
    int m = 0;
    for (;;) {  // sdpa_io.cpp:124-135: skip '*' and '"' comment lines
        if (c.p >= buf.size()) {
            err = "missing mDIM";
            return false;
        }
        size_t e = buf.find('\n', c.p);
        size_t end = (e == std::string::npos) ? buf.size() : e + 1;
        std::string line = buf.substr(c.p, end - c.p);
        c.p = end;
        if (!line.empty() && (line[0] == '*' || line[0] == '"')) continue;
        if (std::sscanf(line.c_str(), "%d", &m) != 1 || m <= 0) {
            err = "bad mDIM";
            return false;
        }
        break;
    }
    int nb = 0;
    if (!c.integer(nb) || nb <= 0) {
        err = "bad nBLOCK";
        return false;
    }
    std::vector<int> given(nb);
    for (int l = 0; l < nb; l++) {
        if (!c.sep() || !c.integer(given[l])) {
            err = "bad bLOCKsTRUCT";
            return false;
        }
    }
    P.m = m;
    if (!P.set_structure(given)) {
        err = "bad bLOCKsTRUCT";
        return false;
    }
    const std::vector<int> &bs = P.bs;
    P.b.assign(m, mk(0.0));
    for (int k = 0; k < m; k++) {
        std::string t;
        if (!c.sep() || !c.token(t)) {
            err = "bad b vector";
            return false;
        }
        P.b[k] = value_of(t);
    }
    struct E {
        int k, l, i, j;
        dd v;
    };
    std::vector<E> ent;
    for (;;) {
        E e;
        std::string t;
        if (!c.sep() || !c.integer(e.k)) break;
        if (!c.sep() || !c.integer(e.l)) break;
        if (!c.sep() || !c.integer(e.i)) break;
        if (!c.sep() || !c.integer(e.j)) break;
        if (!c.sep() || !c.token(t)) break;
        double chk;
        if (std::sscanf(t.c_str(), "%lf", &chk) <= 0) break;
        e.v = value_of(t);
        if (e.k < 0 || e.k > m || e.l < 1 || e.l > nb || e.i < 1 || e.j < 1 || e.i > std::abs(given[e.l - 1]) ||
            e.j > std::abs(given[e.l - 1])) {
            err = "entry out of range in input data";
            return false;
        }
        if (P.gtype[e.l - 1] == 3 && e.i != e.j) {  // sdpa_io.cpp:342-345
            err = "LP entry off the diagonal";
            return false;
        }
        ent.push_back(e);
    }
    const int nsdp = (int)bs.size();
    std::vector<int> cnt((size_t)(m + 1) * nsdp, 0);
    for (auto &e : ent)
        if (P.gtype[e.l - 1] == 1) cnt[(size_t)e.k * nsdp + P.gnum[e.l - 1]]++;
    auto init = [&](Space &sp, int k) {
        for (int l = 0; l < nsdp; l++) {
            int c0 = cnt[(size_t)k * nsdp + l];
            if (!c0) continue;
            sp.index.push_back(l);
            Block b;
            b.n = bs[l];
            b.row.reserve(c0);
            b.col.reserve(c0);
            b.val.reserve(c0);
            sp.blk.push_back(std::move(b));
        }
    };
    init(P.C, 0);
    P.A.assign(m, Space());
    for (int k = 0; k < m; k++) init(P.A[k], k + 1);
    for (auto &e : ent) {
        Space &sp = e.k ? P.A[e.k - 1] : P.C;
        if (P.gtype[e.l - 1] == 3) {
            sp.lp_raw.push_back({P.gnum[e.l - 1] + e.i - 1, e.k ? e.v : dd_neg(e.v)});
            continue;
        }
        Block &b = sp.blk[sp.find(P.gnum[e.l - 1])];
        b.row.push_back(e.i - 1);
        b.col.push_back(e.j - 1);
        b.val.push_back(e.k ? e.v : dd_neg(e.v));  // C = -F0 (sdpa_io.cpp:692)
        b.nnz_count++;
        b.nnz_effect += (e.i == e.j) ? 1 : 2;
    }
    for (auto &b : P.C.blk) b.change_to_dense();
    for (auto &a : P.A)
        for (auto &b : a.blk) b.change_to_dense();
    P.C.finish_lp();
    for (auto &a : P.A) a.finish_lp();
    P.layout();
    P.index();
    P.formula = compute_formula(P, ddr(1.2));
    return true;
}

// sdpa_io.cpp:607-617 (indices sorted, deduplicated) and
// SparseLinearSpace::setElement_LP (sdpa_struct.cpp:1253-1265: a repeated
// index keeps the value set last).
void Space::finish_lp() {
    std::map<int, dd> v;
    for (auto &e : lp_raw) v[e.first] = e.second;
    lp_index.clear();
    lp_val.clear();
    for (auto &e : v) {
        lp_index.push_back(e.first);
        lp_val.push_back(e.second);
    }
    lp_raw.clear();
}

bool HostProblem::set_structure(const std::vector<int> &g) {
    given = g;
    gtype.assign(g.size(), 0);
    gnum.assign(g.size(), 0);
    bs.clear();
    nlp = 0;
    for (size_t l = 0; l < g.size(); l++) {
        if (g[l] >= 2) {
            gtype[l] = 1;
            gnum[l] = (int)bs.size();
            bs.push_back(g[l]);
        } else if (g[l] < 0 || g[l] == 1) {
            gtype[l] = 3;
            gnum[l] = nlp;
            nlp += g[l] < 0 ? -g[l] : 1;
        } else {
            return false;
        }
    }
    return true;
}

void HostProblem::layout() {
    boff.assign(bs.size(), 0);
    nn = 0;
    for (size_t l = 0; l < bs.size(); l++) {
        boff[l] = nn;
        nn += (idx_t)bs[l] * bs[l];
    }
}
void HostProblem::index() {
    cons.assign(bs.size(), {});
    cidx.assign(bs.size(), {});
    for (int i = 0; i < m; i++)
        for (size_t k = 0; k < A[i].index.size(); k++) {
            cons[A[i].index[k]].push_back(i);
            cidx[A[i].index[k]].push_back((int)k);
        }
    // InputData::initialize_index_LP (sdpa_dataset.cpp:395-465)
    lp_cons.assign(nlp, {});
    lp_cidx.assign(nlp, {});
    for (int i = 0; i < m; i++)
        for (size_t k = 0; k < A[i].lp_index.size(); k++) {
            lp_cons[A[i].lp_index[k]].push_back(i);
            lp_cidx[A[i].lp_index[k]].push_back((int)k);
        }
}

// ------------------------------------------------------- initial points --
namespace {
// A value of an initial-point file: the exact pair HI+LO / HI-LO of two
// doubles, or a decimal read as QD reads it.
bool init_value(const std::string &t, dd &v) {
    const char *s = t.c_str();
    char *e1 = nullptr;
    double hi = std::strtod(s, &e1);
    if (e1 != s && (*e1 == '+' || *e1 == '-')) {
        char *e2 = nullptr;
        double lo = std::strtod(e1, &e2);
        if (e2 != e1 && *e2 == '\0') {
            v = dd{hi, lo};
            return true;
        }
    }
    return qd_read(s, v);
}
}  // namespace

void InitPoint::build(const HostProblem &P, double lambda, std::vector<dd> &X, std::vector<dd> &Y,
                      std::vector<dd> &Z) const {
    if (full) {
        X = x;
        Y = y;
        Z = z;
        return;
    }
    // Solutions::initialize (X = Z = lambda I), then the file's settings
    X.assign(P.nt(), mk(0.0));
    for (size_t l = 0; l < P.bs.size(); l++)
        for (int i = 0; i < P.bs[l]; i++) X[P.boff[l] + (idx_t)i * P.bs[l] + i] = mk(lambda);
    for (int i = 0; i < P.nlp; i++) X[P.nn + i] = mk(lambda);
    Z = X;
    Y = y;
    for (const Set &st : sets) {
        std::vector<dd> &v = (st.target == 1) ? Z : X;
        v[st.a] = st.v;
        v[st.b] = st.v;
    }
}

bool read_init(const std::string &path, const HostProblem &P, bool sparse, InitPoint &ip, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open initial-point file " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string buf = ss.str();
    Cursor c(buf);
    ip = InitPoint();
    ip.y.assign(P.m, mk(0.0));
    for (int k = 0; k < P.m; k++) {  // yVec is read with the opposite sign
        std::string t;
        dd v;
        if (!c.sep() || !c.token(t) || !init_value(t, v)) {
            err = "initial point: bad y vector";
            return false;
        }
        ip.y[k] = dd_neg(v);
    }
    const int nB = (int)P.bs.size();
    auto pos = [&](int l, int i, int j) { return P.boff[l] + i + (idx_t)P.bs[l] * j; };
    if (sparse) {
        for (;;) {
            int target, l, i, j;
            std::string t;
            if (!c.sep() || !c.integer(target) || !c.sep() || !c.integer(l) || !c.sep() || !c.integer(i) || !c.sep() ||
                !c.integer(j) || !c.sep() || !c.token(t))
                break;
            double chk;
            if (std::sscanf(t.c_str(), "%lf", &chk) <= 0) break;
            dd v;
            if (!init_value(t, v)) {
                err = "initial point: bad value " + t;
                return false;
            }
            if (l >= 1 && l <= nB) {
                if (i < 1 || j < 1 || i > P.bs[l - 1] || j > P.bs[l - 1]) {
                    err = "initial point: entry out of range";
                    return false;
                }
                ip.sets.push_back({target, pos(l - 1, i - 1, j - 1), pos(l - 1, j - 1, i - 1), v});
            } else {
                if (l < 1 || i != j || i < 1 || i > P.nlp) {
                    err = "initial point: bad LP entry";
                    return false;
                }
                ip.sets.push_back({target, P.nn + i - 1, P.nn + i - 1, v});
            }
        }
    } else {
        for (int pass = 0; pass < 2; pass++) {  // Z, then X
            const int target = pass ? 2 : 1;
            for (int l = 0; l < nB; l++)
                for (int i = 0; i < P.bs[l]; i++)
                    for (int j = 0; j < P.bs[l]; j++) {
                        std::string t;
                        dd v;
                        if (!c.sep() || !c.token(t) || !init_value(t, v)) {
                            err = "initial point: short dense block";
                            return false;
                        }
                        if (i <= j && !dd_is_zero(v)) ip.sets.push_back({target, pos(l, i, j), pos(l, j, i), v});
                    }
            for (int j = 0; j < P.nlp; j++) {
                std::string t;
                dd v;
                if (!c.sep() || !c.token(t) || !init_value(t, v)) {
                    err = "initial point: short LP part";
                    return false;
                }
                if (!dd_is_zero(v)) ip.sets.push_back({target, P.nn + j, P.nn + j, v});
            }
        }
    }
    return true;
}

bool write_point(const std::string &path, const HostProblem &P, const std::vector<dd> &X, const std::vector<dd> &y,
                 const std::vector<dd> &Z, std::string &err) {
    if ((idx_t)X.size() != P.nt() || (idx_t)Z.size() != P.nt() || (int)y.size() != P.m) {
        err = "write_point: no solution of this problem";
        return false;
    }
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) {
        err = "cannot write " + path;
        return false;
    }
    auto val = [&](dd v) {  // exact: HI then the signed LO, each a round-trip double
        std::fprintf(fp, "%.17e%+.17e", v.hi, v.lo);
    };
    for (int k = 0; k < P.m; k++) {  // the file's y is the solver's negated
        std::fputc(' ', fp);  // the readers skip one or more separators before each value
        val(dd_neg(y[k]));
    }
    std::fputc('\n', fp);
    const int nB = (int)P.bs.size();
    for (int target = 1; target <= 2; target++) {
        const std::vector<dd> &v = (target == 1) ? Z : X;
        for (int l = 0; l < nB; l++) {
            const int n = P.bs[l];
            for (int i = 0; i < n; i++)
                for (int j = i; j < n; j++) {
                    dd e = v[P.boff[l] + i + (idx_t)n * j];
                    if (i != j && dd_is_zero(e)) continue;
                    std::fprintf(fp, "%d %d %d %d ", target, l + 1, i + 1, j + 1);
                    val(e);
                    std::fputc('\n', fp);
                }
        }
        for (int i = 0; i < P.nlp; i++) {
            std::fprintf(fp, "%d %d %d %d ", target, nB + 1, i + 1, i + 1);
            val(v[P.nn + i]);
            std::fputc('\n', fp);
        }
    }
    bool ok = std::fclose(fp) == 0;
    if (!ok) err = "write failed: " + path;
    return ok;
}

bool read_params(const std::string &path, Params &p) {
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return false;
    int r = 0;
    r += std::fscanf(fp, "%d%*[^\n]", &p.maxIteration);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.epsilonStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.lambdaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.omegaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.lowerBound);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.upperBound);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.betaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.betaBar);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.gammaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &p.epsilonDash);
    std::fclose(fp);
    return r == 10;
}

// Newton::computeFormula_SDP (sdpa_newton.cpp:694-819), evaluated in dd.
std::vector<int> compute_formula(const HostProblem &P, ddr Kappa) {
    const int m = P.m, nB = (int)P.bs.size();
    std::vector<int> up((size_t)m * nB, 0), f((size_t)m * nB, 0);
    for (int l = 0; l < nB; l++) {
        const auto &cs = P.cons[l];
        const auto &ci = P.cidx[l];
        for (size_t k1 = 0; k1 < cs.size(); k1++) {
            int i = cs[k1], inz = P.A[i].blk[ci[k1]].nnz_effect, u = inz;
            for (size_t k2 = 0; k2 < cs.size(); k2++) {
                int j = cs[k2], jnz = P.A[j].blk[ci[k2]].nnz_effect;
                if (jnz < inz || (jnz == inz && j < i)) u += jnz;
            }
            up[(size_t)i * nB + l] = u;
        }
    }
    for (int l = 0; l < nB; l++) {
        const auto &cs = P.cons[l];
        const auto &ci = P.cidx[l];
        for (size_t k = 0; k < cs.size(); k++) {
            int i = cs[k];
            const Block &Ai = P.A[i].blk[ci[k]];
            ddr inz = Ai.nnz_effect, n = Ai.n, u = up[(size_t)i * nB + l];
            ddr f1 = Kappa * n * inz + n * n * n + Kappa * u;
            ddr f2 = Kappa * n * inz + Kappa * (n + 1) * u;
            ddr f3 = Kappa * (2 * Kappa * inz + 1) * u / Kappa;
            int r;
            if (Ai.dense)
                r = (f1 < f2) ? 0 : 1;
            else if (f1 < f2 && f1 < f3)
                r = 0;
            else if (f2 < f3)
                r = 1;
            else
                r = 2;
            f[(size_t)i * nB + l] = r;
        }
    }
    return f;
}

// Algorithmic dd operations of one Schur assembly (compute_bMat_dense_SDP,
// sdpa_newton.cpp:979-1068, and compute_bMat_dense_LP, :1162-1197): per row
// i of block l, A_i Z^{-1} costs 2 nnz_i n (sparse) or 2 n^3 (dense), F1's
// X (A_i Z^{-1}) another 2 n^3; per pair (i, j) the row visits, F1 costs
// 2 nnz_j (the inner product with A_j), F2 2 n nnz_j for a sparse A_j
// (calF2's strided dots) and 2 n^2 (+ 2 n^3 once per row for X F) for a
// dense one, F3 3 nnz_i nnz_j (calF3: two products and a sum per entry
// pair); each LP entry a pair shares 3.  nnz = NonZeroEffect.
double schur_ddflops(const HostProblem &P) {
    const int nB = (int)P.bs.size();
    double total = 0.0;
    for (int l = 0; l < nB; l++) {
        const auto &cs = P.cons[l];
        const auto &ci = P.cidx[l];
        const double n = P.bs[l];
        for (size_t k1 = 0; k1 < cs.size(); k1++) {
            const int i = cs[k1];
            const Block &Ai = P.A[i].blk[ci[k1]];
            const int inz = Ai.nnz_effect;
            const int f = P.formula[(size_t)i * nB + l];
            const double aiz = Ai.dense ? 2.0 * n * n * n : 2.0 * inz * n;
            if (f == 0) total += aiz + 2.0 * n * n * n;
            if (f == 1) total += aiz;
            bool g = false;
            for (size_t k2 = 0; k2 < cs.size(); k2++) {
                const int j = cs[k2];
                const Block &Aj = P.A[j].blk[ci[k2]];
                const int jnz = Aj.nnz_effect;
                if (inz < jnz || (inz == jnz && i < j)) continue;
                if (f == 0)
                    total += 2.0 * (Aj.dense ? n * n : jnz);
                else if (f == 1) {
                    if (Aj.dense) {
                        if (!g) total += 2.0 * n * n * n;
                        g = true;
                        total += 2.0 * n * n;
                    } else {
                        total += 2.0 * n * jnz;
                    }
                } else {
                    total += 3.0 * (double)inz * (double)jnz;
                }
            }
        }
    }
    for (size_t l = 0; l < P.lp_cons.size(); l++) {
        const double c = (double)P.lp_cons[l].size();
        total += 3.0 * c * (c + 1) / 2;
    }
    return total;
}

// ----------------------------------------------------------------- upload --
template <typename T>
static T *dev_copy(DeviceProblem &dp, const std::vector<T> &h, hipStream_t s) {
    if (h.empty()) return nullptr;
    T *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(T) * h.size()) != hipSuccess) return nullptr;
    dp.allocs.push_back(d);
    (void)s;
    if (hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
}
template <typename T>
static T *dev_alloc(DeviceProblem &dp, size_t n) {
    if (n == 0) return nullptr;
    T *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(T) * n) != hipSuccess) return nullptr;
    dp.allocs.push_back(d);
    return d;
}

int upload_problem(const HostProblem &P, DeviceProblem &dp, hipStream_t s, std::string &err) {
    const int m = P.m, nB = (int)P.bs.size();
    dp.m = m;
    dp.bs = P.bs;
    dp.boff = P.boff;
    dp.nn = P.nn;
    dp.nlp = P.nlp;
    dp.d_b = dev_copy(dp, P.b, s);
    // Shared sparse-entry and dense-value pools.
    std::vector<int> sp_row, sp_col;
    std::vector<dd> sp_val, dense;
    auto add_block = [&](const Block &b, int &type, int &nnz, idx_t &off) {
        if (b.dense) {
            type = 1;
            nnz = b.n * b.n;
            off = (idx_t)dense.size();
            dense.insert(dense.end(), b.dval.begin(), b.dval.end());
        } else {
            type = 0;
            nnz = b.nnz_count;
            off = (idx_t)sp_row.size();
            sp_row.insert(sp_row.end(), b.row.begin(), b.row.end());
            sp_col.insert(sp_col.end(), b.col.begin(), b.col.end());
            sp_val.insert(sp_val.end(), b.val.begin(), b.val.end());
        }
    };
    // Inner-product tasks: constraint i, its blocks in SDP_sp_index order.
    std::vector<idx_t> tptr(m + 1, 0);
    std::vector<int> ttype, tblock, tnnz;
    std::vector<idx_t> toff;
    // Remember each (i, block-slot) pool location for the Schur data.
    std::vector<std::vector<std::pair<int, idx_t>>> loc(m);
    for (int i = 0; i < m; i++) {
        tptr[i] = (idx_t)ttype.size();
        for (size_t k = 0; k < P.A[i].index.size(); k++) {
            int type, nnz;
            idx_t off;
            add_block(P.A[i].blk[k], type, nnz, off);
            ttype.push_back(type);
            tblock.push_back(P.A[i].index[k]);
            tnnz.push_back(nnz);
            toff.push_back(off);
            loc[i].push_back({type, off});
        }
    }
    tptr[m] = (idx_t)ttype.size();
    std::vector<idx_t> cptr(2, 0);
    std::vector<int> cty, cbl, cnz;
    std::vector<idx_t> coff;
    for (size_t k = 0; k < P.C.index.size(); k++) {
        int type, nnz;
        idx_t off;
        add_block(P.C.blk[k], type, nnz, off);
        cty.push_back(type);
        cbl.push_back(P.C.index[k]);
        cnz.push_back(nnz);
        coff.push_back(off);
    }
    cptr[1] = (idx_t)cty.size();
    dp.ntask = (int)ttype.size();
    dp.h_task_type = ttype;
    dp.d_task_ptr = dev_copy(dp, tptr, s);
    dp.d_task_type = dev_copy(dp, ttype, s);
    dp.d_task_block = dev_copy(dp, tblock, s);
    dp.d_task_nnz = dev_copy(dp, tnnz, s);
    dp.d_task_off = dev_copy(dp, toff, s);
    dp.nctask = (int)cty.size();
    dp.h_ctask_type = cty;
    dp.d_ctask_ptr = dev_copy(dp, cptr, s);
    dp.d_ctask_type = dev_copy(dp, cty, s);
    dp.d_ctask_block = dev_copy(dp, cbl, s);
    dp.d_ctask_nnz = dev_copy(dp, cnz, s);
    dp.d_ctask_off = dev_copy(dp, coff, s);
    dp.d_sp_row = dev_copy(dp, sp_row, s);
    dp.d_sp_col = dev_copy(dp, sp_col, s);
    dp.d_sp_val = dev_copy(dp, sp_val, s);
    dp.d_dense = dev_copy(dp, dense, s);

    // LP entries of each A_i and of C for the inner products.
    {
        std::vector<idx_t> lptr(m + 1, 0), cptr2(2, 0);
        std::vector<int> lidx, clidx;
        std::vector<dd> lval, clval;
        for (int i = 0; i < m; i++) {
            lptr[i] = (idx_t)lidx.size();
            lidx.insert(lidx.end(), P.A[i].lp_index.begin(), P.A[i].lp_index.end());
            lval.insert(lval.end(), P.A[i].lp_val.begin(), P.A[i].lp_val.end());
        }
        lptr[m] = (idx_t)lidx.size();
        clidx = P.C.lp_index;
        clval = P.C.lp_val;
        cptr2[1] = (idx_t)clidx.size();
        dp.d_lp_ptr = dev_copy(dp, lptr, s);
        dp.d_lp_idx = dev_copy(dp, lidx, s);
        dp.d_lp_val = dev_copy(dp, lval, s);
        dp.d_clp_ptr = dev_copy(dp, cptr2, s);
        dp.d_clp_idx = dev_copy(dp, clidx, s);
        dp.d_clp_val = dev_copy(dp, clval, s);
    }
    // LP part of the Schur matrix (Newton::compute_bMat_dense_LP,
    // sdpa_newton.cpp:1162-1197): LP entry l adds x_l z_l^-1 a_il a_jl to B_ij
    // for every pair k1 <= k2 of its constraint list (constraints ascending),
    // the entries in ascending l.  Grouped here per pair (i, j), i <= j.
    {
        std::map<std::pair<int, int>, std::vector<int>> pl;  // pair -> (l, k1, k2) triples
        for (int l = 0; l < P.nlp; l++) {
            const auto &cs = P.lp_cons[l];
            for (size_t k1 = 0; k1 < cs.size(); k1++)
                for (size_t k2 = k1; k2 < cs.size(); k2++) {
                    auto &v = pl[{cs[k1], cs[k2]}];
                    v.push_back(l);
                    v.push_back((int)k1);
                    v.push_back((int)k2);
                }
        }
        std::vector<int> pi, pj, ll;
        std::vector<idx_t> ptr;
        std::vector<dd> ai, aj;
        for (auto &e : pl) {
            pi.push_back(e.first.first);
            pj.push_back(e.first.second);
            ptr.push_back((idx_t)ll.size());
            for (size_t q = 0; q < e.second.size(); q += 3) {
                int l = e.second[q], k1 = e.second[q + 1], k2 = e.second[q + 2];
                int i = P.lp_cons[l][k1], j = P.lp_cons[l][k2];
                ll.push_back(l);
                ai.push_back(P.A[i].lp_val[P.lp_cidx[l][k1]]);
                aj.push_back(P.A[j].lp_val[P.lp_cidx[l][k2]]);
            }
        }
        ptr.push_back((idx_t)ll.size());
        dp.n_lp_pairs = (int)pi.size();
        dp.d_lpp_i = dev_copy(dp, pi, s);
        dp.d_lpp_j = dev_copy(dp, pj, s);
        dp.d_lpp_ptr = dev_copy(dp, ptr, s);
        dp.d_lpp_l = dev_copy(dp, ll, s);
        dp.d_lpp_ai = dev_copy(dp, ai, s);
        dp.d_lpp_aj = dev_copy(dp, aj, s);
    }

    // Element update lists for sum_i y_i A_i (multi_plusToA order: i
    // ascending, entries in file order, off-diagonals to both positions;
    // LP entries at nn + index).
    {
        const idx_t nt = P.nt();
        std::vector<std::vector<std::pair<int, dd>>> el(nt);
        for (int i = 0; i < m; i++)
            for (size_t k = 0; k < P.A[i].index.size(); k++) {
                int l = P.A[i].index[k], n = P.bs[l];
                idx_t o = P.boff[l];
                const Block &b = P.A[i].blk[k];
                if (b.dense) {
                    for (idx_t e = 0; e < (idx_t)n * n; e++)
                        if (!dd_is_zero(b.dval[e])) el[o + e].push_back({i, b.dval[e]});
                } else {
                    for (int q = 0; q < b.nnz_count; q++) {
                        int r = b.row[q], cc = b.col[q];
                        el[o + r + (idx_t)n * cc].push_back({i, b.val[q]});
                        if (r != cc) el[o + cc + (idx_t)n * r].push_back({i, b.val[q]});
                    }
                }
            }
        for (int i = 0; i < m; i++)
            for (size_t k = 0; k < P.A[i].lp_index.size(); k++)
                el[P.nn + P.A[i].lp_index[k]].push_back({i, P.A[i].lp_val[k]});
        std::vector<idx_t> ptr(nt + 1, 0);
        std::vector<int> idx;
        std::vector<dd> val;
        for (idx_t e = 0; e < nt; e++) {
            ptr[e] = (idx_t)idx.size();
            for (auto &pr : el[e]) {
                idx.push_back(pr.first);
                val.push_back(pr.second);
            }
        }
        ptr[nt] = (idx_t)idx.size();
        dp.d_el_ptr = dev_copy(dp, ptr, s);
        dp.d_el_idx = dev_copy(dp, idx, s);
        dp.d_el_val = dev_copy(dp, val, s);
        std::vector<std::vector<dd>> cel(nt);
        for (size_t k = 0; k < P.C.index.size(); k++) {
            int l = P.C.index[k], n = P.bs[l];
            idx_t o = P.boff[l];
            const Block &b = P.C.blk[k];
            if (b.dense) {
                for (idx_t e = 0; e < (idx_t)n * n; e++)
                    if (!dd_is_zero(b.dval[e])) cel[o + e].push_back(b.dval[e]);
            } else {
                for (int q = 0; q < b.nnz_count; q++) {
                    int r = b.row[q], cc = b.col[q];
                    cel[o + r + (idx_t)n * cc].push_back(b.val[q]);
                    if (r != cc) cel[o + cc + (idx_t)n * r].push_back(b.val[q]);
                }
            }
        }
        for (size_t k = 0; k < P.C.lp_index.size(); k++) cel[P.nn + P.C.lp_index[k]].push_back(P.C.lp_val[k]);
        std::vector<idx_t> cp(nt + 1, 0);
        std::vector<dd> cv;
        for (idx_t e = 0; e < nt; e++) {
            cp[e] = (idx_t)cv.size();
            for (auto &v : cel[e]) cv.push_back(v);
        }
        cp[nt] = (idx_t)cv.size();
        dp.d_cel_ptr = dev_copy(dp, cp, s);
        dp.d_cel_val = dev_copy(dp, cv, s);
    }

    // Dense inner-product tasks (A_i then C), y offset = block offset.
    {
        std::vector<sdk::DotTaskOff> dt, dct;
        dp.d_task_val = dev_alloc<dd>(dp, std::max(1, dp.ntask));
        dp.d_ctask_val = dev_alloc<dd>(dp, std::max(1, dp.nctask));
        dd *dense_base = dp.d_dense;
        for (int t = 0; t < dp.ntask; t++)
            if (ttype[t] == 1) dt.push_back({dense_base + toff[t], P.boff[tblock[t]], (idx_t)tnnz[t], dp.d_task_val + t});
        for (int t = 0; t < dp.nctask; t++)
            if (cty[t] == 1) dct.push_back({dense_base + coff[t], P.boff[cbl[t]], (idx_t)cnz[t], dp.d_ctask_val + t});
        dp.n_dense_tasks = (int)dt.size();
        dp.n_dense_ctasks = (int)dct.size();
        for (auto &t : dt) dp.dense_task_maxn = std::max(dp.dense_task_maxn, t.n);
        for (auto &t : dct) dp.dense_ctask_maxn = std::max(dp.dense_ctask_maxn, t.n);
        dp.d_dense_tasks = dev_copy(dp, dt, s);
        dp.d_dense_ctasks = dev_copy(dp, dct, s);
        dp.d_boff = dev_copy(dp, P.boff, s);
        dp.d_bsize = dev_copy(dp, P.bs, s);
    }

    // Schur data per block.
    size_t work_need = 0;
    int max_pairs = 0;
    dp.schur.assign(nB, SchurBlock());
    struct PendingTasks {
        int l;
        std::vector<int> psi, psj;
    };
    std::vector<PendingTasks> pending;
    std::vector<std::vector<int>> wslots(nB);
    std::vector<std::vector<idx_t>> offs(nB);
    for (int l = 0; l < nB; l++) {
        SchurBlock &S = dp.schur[l];
        const int n = P.bs[l];
        S.n = n;
        const auto &cs = P.cons[l];
        const auto &ci = P.cidx[l];
        const int ns = (int)cs.size();
        std::vector<int> type(ns), nnz(ns), eff(ns);
        std::vector<idx_t> off(ns);
        std::vector<int> wslot(ns, -1);
        int nw = 0;
        for (int k = 0; k < ns; k++) {
            int i = cs[k];
            const Block &b = P.A[i].blk[ci[k]];
            type[k] = b.dense ? 1 : 0;
            nnz[k] = b.nnz_count;
            eff[k] = b.nnz_effect;
            off[k] = loc[i][ci[k]].second;
            int fm = P.formula[(size_t)i * nB + l];
            S.ci.push_back(i);
            S.form.push_back(fm);
            if (fm == 0)
                S.slots_f1.push_back(k);
            else if (fm == 1)
                S.slots_f2.push_back(k);
            else
                S.slots_f3.push_back(k);
            if (fm != 2) {
                wslot[k] = nw++;
                (type[k] ? S.de_slots : S.sp_slots).push_back(k);
            }
        }
        // pairs in the order of compute_bMat_dense_SDP's loops (sdpa_newton.cpp:989-1026)
        std::vector<int> pi, pj, psi, psj, pf;
        std::vector<int> need_g(ns, 0);
        for (int k1 = 0; k1 < ns; k1++)
            for (int k2 = 0; k2 < ns; k2++) {
                int i = cs[k1], j = cs[k2];
                if (eff[k1] < eff[k2] || (eff[k1] == eff[k2] && i < j)) continue;
                pi.push_back(i);
                pj.push_back(j);
                psi.push_back(k1);
                psj.push_back(k2);
                pf.push_back(S.form[k1]);
                if (S.form[k1] == 0 || (S.form[k1] == 1 && type[k2] == 1)) need_g[k1] = 1;
            }
        for (int k = 0; k < ns; k++)
            if (need_g[k] && S.form[k] != 2) S.g_slots.push_back(k);
        S.npairs = (int)pi.size();
        max_pairs = std::max(max_pairs, S.npairs);
        S.h_pair_form = pf;
        {
            std::vector<int> lf2, lot;
            for (int q = 0; q < S.npairs; q++) {
                if (type[psj[q]] != 0) continue;  // dense A_j: dot tasks
                (pf[q] == 1 ? lf2 : lot).push_back(q);
            }
            S.n_f2 = (int)lf2.size();
            S.n_other = (int)lot.size();
            S.d_list_f2 = dev_copy(dp, lf2, s);
            S.d_list_other = dev_copy(dp, lot, s);
        }
        S.d_pair_i = dev_copy(dp, pi, s);
        S.d_pair_j = dev_copy(dp, pj, s);
        S.d_pair_slot_i = dev_copy(dp, psi, s);
        S.d_pair_slot_j = dev_copy(dp, psj, s);
        S.d_pair_form = dev_copy(dp, pf, s);
        S.d_type = dev_copy(dp, type, s);
        S.d_nnz = dev_copy(dp, nnz, s);
        S.d_off = dev_copy(dp, off, s);
        S.d_wslot = dev_copy(dp, wslot, s);
        S.d_sp_slots = dev_copy(dp, S.sp_slots, s);
        S.d_de_slots = dev_copy(dp, S.de_slots, s);
        S.d_g_slots = dev_copy(dp, S.g_slots, s);
        pending.push_back({l, psi, psj});
        wslots[l] = wslot;
        offs[l] = off;
        // row lists of A_i * Zinv for sparse F1/F2 constraints (Raxpy order)
        std::vector<idx_t> rptr((size_t)ns * (n + 1), 0);
        std::vector<int> rsrc;
        std::vector<dd> rval;
        for (int k = 0; k < ns; k++) {
            const Block &b = P.A[cs[k]].blk[ci[k]];
            std::vector<std::vector<std::pair<int, dd>>> rows(n);
            if (!b.dense && S.form[k] != 2) {
                for (int q = 0; q < b.nnz_count; q++) {
                    int r = b.row[q], cc = b.col[q];
                    if (r != cc) {
                        rows[r].push_back({cc, b.val[q]});
                        rows[cc].push_back({r, b.val[q]});
                    } else {
                        rows[cc].push_back({cc, b.val[q]});
                    }
                }
            }
            for (int r = 0; r < n; r++) {
                rptr[(size_t)k * (n + 1) + r] = (idx_t)rsrc.size();
                for (auto &pr : rows[r]) {
                    rsrc.push_back(pr.first);
                    rval.push_back(pr.second);
                }
            }
            rptr[(size_t)k * (n + 1) + n] = (idx_t)rsrc.size();
        }
        S.d_rl_ptr = dev_copy(dp, rptr, s);
        S.d_rl_src = dev_copy(dp, rsrc, s);
        S.d_rl_val = dev_copy(dp, rval, s);
        work_need = std::max(work_need, (size_t)nw * n * n);
    }
    dp.work_elems = work_need;
    dp.d_work1 = dev_alloc<dd>(dp, work_need);
    dp.d_work2 = dev_alloc<dd>(dp, work_need);
    dp.d_pair_val = dev_alloc<dd>(dp, std::max(1, max_pairs));
    // One SDP block: every element of B takes at most one pair value, so B
    // is gathered element by element (coalesced) instead of zeroed and
    // scattered pair by pair (half the scatter's writes stride by m).
    if (dp.schur.size() == 1 && dp.schur[0].npairs > 0 && m <= 16384) {
        std::vector<int> pidx((size_t)m * m, -1);
        std::vector<int> hpi, hpj;
        {
            const auto &cs = P.cons[0];
            std::vector<int> eff(cs.size());
            for (size_t k = 0; k < cs.size(); k++) eff[k] = P.A[cs[k]].blk[P.cidx[0][k]].nnz_effect;
            int p = 0;
            for (size_t k1 = 0; k1 < cs.size(); k1++)
                for (size_t k2 = 0; k2 < cs.size(); k2++) {
                    const int i = cs[k1], j = cs[k2];
                    if (eff[k1] < eff[k2] || (eff[k1] == eff[k2] && i < j)) continue;
                    pidx[i + (size_t)m * j] = p;
                    pidx[j + (size_t)m * i] = p;
                    p++;
                }
            if (p != dp.schur[0].npairs) {
                err = "Schur pair map: pair count mismatch";
                return -1;
            }
        }
        dp.d_pidx = dev_copy(dp, pidx, s);
    }
    // Dot tasks for pairs with a dense A_j, now that work2 has an address.
    for (auto &pt : pending) {
        SchurBlock &S = dp.schur[pt.l];
        const idx_t nsq = (idx_t)S.n * S.n;
        const auto &cs = P.cons[pt.l];
        const auto &ci = P.cidx[pt.l];
        std::vector<sdk::DotTask> tasks;
        for (int p = 0; p < S.npairs; p++) {
            int sj = pt.psj[p], si = pt.psi[p];
            if (!P.A[cs[sj]].blk[ci[sj]].dense) continue;
            tasks.push_back({dp.d_dense + offs[pt.l][sj], dp.d_work2 + (idx_t)wslots[pt.l][si] * nsq, nsq, dp.d_pair_val + p});
        }
        S.ndense_pairs = (int)tasks.size();
        S.d_pair_tasks = dev_copy(dp, tasks, s);
    }
    size_t maxslots = 1;
    for (auto &S : dp.schur) maxslots = std::max(maxslots, std::max(S.de_slots.size(), S.g_slots.size()));
    dp.ptr_cap = maxslots;
    dp.d_ptrs = dev_alloc<char>(dp, 3 * sizeof(void *) * maxslots);
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        err = hipGetErrorString(e);
        return -1;
    }
    return 0;
}

void free_problem(DeviceProblem &dp) {
    for (void *p : dp.allocs) hipFree(p);
    dp.allocs.clear();
    if (dp.d_dot_partial) (void)hipFree(dp.d_dot_partial);
    dp.d_dot_partial = nullptr;
    dp.dot_partial_cap = 0;
}

// ------------------------------------------------------- inner products --
// One thread per sparse task: the 4-way grouped sum of getInnerProduct
// (sdpa_linear.cpp:202-290); dense tasks are left to the chunked dot.
__device__ dd sparse_inner(const int *row, const int *col, const dd *val, int nnz, const dd *W, int n) {
    auto term = [&](int q) {
        int i = row[q], j = col[q];
        if (i == j) return dd_mul(val[q], W[i + (idx_t)n * j]);
        return dd_mul(val[q], dd_add(W[i + (idx_t)n * j], W[j + (idx_t)n * i]));
    };
    dd ret = mk(0.0);
    int amari = nnz % 4, shou = nnz / 4;
    for (int q = 0; q < amari; q++) ret = dd_add_eq(ret, term(q));
    for (int c = 0, q = amari; c < shou; c++, q += 4) {
        dd r1 = term(q), r2 = term(q + 1), r3 = term(q + 2), r4 = term(q + 3);
        ret = dd_add_eq(ret, dd_add(dd_add(dd_add(r1, r2), r3), r4));
    }
    return ret;
}

// One wave per task.  The terms (and the four-term groups of
// getInnerProduct's unrolled loop) are independent, so the lanes form them;
// lane 0 then adds the leading single terms and the group sums in order.
__global__ __launch_bounds__(64) void k_sparse_tasks(int ntask, const int *type, const int *block, const int *nnz,
                                                     const idx_t *off, const int *row, const int *col, const dd *val,
                                                     const idx_t *boff, const int *bsize, const dd *W, dd *out) {
    __shared__ dd g[256];
    const int t = blockIdx.x, lane = threadIdx.x;
    if (type[t] != 0) return;
    const int l = block[t], nz = nnz[t], n = bsize[l];
    const idx_t o = off[t];
    const int *rw = row + o, *cl = col + o;
    const dd *vl = val + o, *Wl = W + boff[l];
    auto term = [&](int q) {
        int i = rw[q], j = cl[q];
        if (i == j) return dd_mul(vl[q], Wl[i + (idx_t)n * j]);
        return dd_mul(vl[q], dd_add(Wl[i + (idx_t)n * j], Wl[j + (idx_t)n * i]));
    };
    const int amari = nz % 4, shou = nz / 4, nitems = amari + shou;
    dd ret = mk(0.0);
    for (int base = 0; base < nitems; base += 256) {
        const int cnt = nitems - base < 256 ? nitems - base : 256;
        for (int k = lane; k < cnt; k += 64) {
            const int it = base + k;
            if (it < amari) {
                g[k] = term(it);
            } else {
                const int q = amari + 4 * (it - amari);
                g[k] = dd_add(dd_add(dd_add(term(q), term(q + 1)), term(q + 2)), term(q + 3));
            }
        }
        __syncthreads();
        if (lane == 0)
            for (int k = 0; k < cnt; k++) ret = dd_add_eq(ret, g[k]);
        __syncthreads();
    }
    if (lane == 0) out[t] = ret;
}
// out[i] = 0 + sum of its task values in block order, then its LP terms
// a_l w_l in ascending l (Lal::getInnerProduct, sdpa_linear.cpp:1616-1650)
__global__ void k_combine_tasks(int m, const idx_t *ptr, const dd *tv, const idx_t *lptr, const int *lidx,
                                const dd *lval, const dd *Wlp, dd *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    dd r = mk(0.0);
    for (idx_t t = ptr[i]; t < ptr[i + 1]; t++) r = dd_add_eq(r, tv[t]);
    for (idx_t q = lptr[i]; q < lptr[i + 1]; q++) r = dd_add_eq(r, dd_mul(lval[q], Wlp[lidx[q]]));
    out[i] = r;
}

// Dense tasks: one thread per chunk while chunks are short; one workgroup
// per chunk (coalesced loads) once they are long.  Same sums either way.
static void dense_dots(hipStream_t s, DeviceProblem &dp, const sdk::DotTaskOff *tasks, int ntasks, idx_t maxn,
                       const dd *W, int T, bool fast) {
    if (ntasks <= 0) return;
    const size_t need = (size_t)T * ntasks;
    bool chunked = maxn / T > 64;
    if (chunked && need > dp.dot_partial_cap) {
        if (dp.d_dot_partial) (void)hipFree(dp.d_dot_partial);
        dp.d_dot_partial = nullptr;
        dp.dot_partial_cap = 0;
        if (hipMalloc((void **)&dp.d_dot_partial, sizeof(dd) * need) == hipSuccess)
            dp.dot_partial_cap = need;
        else
            chunked = false;
    }
    if (chunked && fast)
        sdk::dot_tasks_off_fast(s, tasks, ntasks, W, T, dp.d_dot_partial);
    else if (chunked)
        sdk::dot_tasks_off(s, tasks, ntasks, W, T, dp.d_dot_partial);
    else
        sdk::dot_batch_off(s, tasks, ntasks, W, T);
}

void inner_A(hipStream_t s, DeviceProblem &dp, const dd *W, dd *out, int T, bool fast) {
    if (dp.ntask > 0)
        k_sparse_tasks<<<dp.ntask, 64, 0, s>>>(dp.ntask, dp.d_task_type, dp.d_task_block, dp.d_task_nnz,
                                                              dp.d_task_off, dp.d_sp_row, dp.d_sp_col, dp.d_sp_val,
                                                              dp.d_boff, dp.d_bsize, W, dp.d_task_val);
    dense_dots(s, dp, dp.d_dense_tasks, dp.n_dense_tasks, dp.dense_task_maxn, W, T, fast);
    k_combine_tasks<<<(dp.m + 255) / 256, 256, 0, s>>>(dp.m, dp.d_task_ptr, dp.d_task_val, dp.d_lp_ptr, dp.d_lp_idx,
                                                       dp.d_lp_val, W + dp.nn, out);
}
void inner_C(hipStream_t s, DeviceProblem &dp, const dd *W, dd *out, int T, bool fast) {
    if (dp.nctask > 0)
        k_sparse_tasks<<<dp.nctask, 64, 0, s>>>(dp.nctask, dp.d_ctask_type, dp.d_ctask_block,
                                                               dp.d_ctask_nnz, dp.d_ctask_off, dp.d_sp_row, dp.d_sp_col,
                                                               dp.d_sp_val, dp.d_boff, dp.d_bsize, W, dp.d_ctask_val);
    dense_dots(s, dp, dp.d_dense_ctasks, dp.n_dense_ctasks, dp.dense_ctask_maxn, W, T, fast);
    k_combine_tasks<<<1, 32, 0, s>>>(1, dp.d_ctask_ptr, dp.d_ctask_val, dp.d_clp_ptr, dp.d_clp_idx, dp.d_clp_val,
                                     W + dp.nn, out);
}

// ------------------------------------------------------------ sum y_i A_i --
// SDP elements accumulate with +=; LP elements (e >= nn) with r = r + a y
// (Lal::plus's LP loop, sdpa_linear.cpp:1764-1771), a different rounding.
__global__ void k_plus_A(idx_t nt, idx_t nn, const idx_t *ptr, const int *idx, const dd *val, const dd *y, dd *R) {
    idx_t e = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    dd acc = mk(0.0);
    if (e < nn)
        for (idx_t q = ptr[e]; q < ptr[e + 1]; q++) acc = dd_add_eq(acc, dd_mul(val[q], y[idx[q]]));
    else
        for (idx_t q = ptr[e]; q < ptr[e + 1]; q++) acc = dd_add(acc, dd_mul(val[q], y[idx[q]]));
    R[e] = acc;
}
__global__ void k_plus_C(idx_t nt, idx_t nn, const idx_t *ptr, const dd *val, dd *R) {
    idx_t e = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nt) return;
    idx_t q0 = ptr[e], q1 = ptr[e + 1];
    if (q0 == q1) return;
    dd acc = R[e];
    if (e < nn)
        for (idx_t q = q0; q < q1; q++) acc = dd_add_eq(acc, val[q]);
    else
        for (idx_t q = q0; q < q1; q++) acc = dd_add(acc, val[q]);
    R[e] = acc;
}
void plus_A(hipStream_t s, DeviceProblem &dp, const dd *y, dd *R) {
    const idx_t nt = dp.nn + dp.nlp;
    k_plus_A<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(nt, dp.nn, dp.d_el_ptr, dp.d_el_idx, dp.d_el_val, y, R);
}
void plus_C(hipStream_t s, DeviceProblem &dp, dd *R) {
    const idx_t nt = dp.nn + dp.nlp;
    k_plus_C<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(nt, dp.nn, dp.d_cel_ptr, dp.d_cel_val, R);
}

// ------------------------------------------------------------------ Schur --
// work1[w] = A_i * Zinv for sparse A_i via row lists: row r accumulates
// v * Zinv[src, :] in Raxpy order (sdpa_linear.cpp:843-866).
__global__ void k_rowlist_mul(int n, int ns, const int *slot_list, int nslots, const int *wslot, const idx_t *rl_ptr,
                              const int *rl_src, const dd *rl_val, const dd *Zinv, dd *work1) {
    idx_t e = (idx_t)blockIdx.x * blockDim.x + threadIdx.x;
    int which = blockIdx.y;
    if (which >= nslots || e >= (idx_t)n * n) return;
    int k = slot_list[which];
    int r = (int)(e % n), t = (int)(e / n);
    const idx_t *p = rl_ptr + (idx_t)k * (n + 1);
    dd acc = mk(0.0);
    for (idx_t q = p[r]; q < p[r + 1]; q++) {
        dd v = rl_val[q];
        if (dd_is_zero(v)) continue;  // Raxpy returns when da == 0
        acc = dd_add_eq(acc, dd_mul(v, Zinv[rl_src[q] + (idx_t)n * t]));
    }
    work1[(idx_t)wslot[k] * n * n + e] = acc;
}

// Pair values computed by one thread: F1 with sparse A_j, F2 with sparse
// A_j, and F3 (sdpa_newton.cpp:895-977).
__global__ void k_pair_thread(int nlist, const int *list, int n, const int *pform, const int *psi, const int *psj,
                              const int *type, const int *nnz, const idx_t *off, const int *wslot, const int *row,
                              const int *col, const dd *val, const dd *X, const dd *Zinv, const dd *work1,
                              const dd *work2, dd *pv) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlist) return;
    const int p = list[t];
    int f = pform[p], si = psi[p], sj = psj[p];
    if (type[sj] != 0) return;  // dense A_j: chunked dot kernel
    const int *rj = row + off[sj], *cj = col + off[sj];
    const dd *vj = val + off[sj];
    const int nj = nnz[sj];
    const idx_t nsq = (idx_t)n * n;
    dd ret = mk(0.0);
    if (f == 0) {
        ret = sparse_inner(rj, cj, vj, nj, work2 + (idx_t)wslot[si] * nsq, n);
    } else if (f == 1) {
        const dd *F = work1 + (idx_t)wslot[si] * nsq;
        for (int q = 0; q < nj; q++) {
            int a = rj[q], b = cj[q];
            dd v1 = vj[q];
            dd v2 = mk(0.0);
            for (int t = 0; t < n; t++) v2 = dd_add_eq(v2, dd_mul(X[a + (idx_t)n * t], F[t + (idx_t)n * b]));
            ret = dd_add_eq(ret, dd_mul(v1, v2));
            if (a != b) {
                v2 = mk(0.0);
                for (int t = 0; t < n; t++) v2 = dd_add_eq(v2, dd_mul(X[b + (idx_t)n * t], F[t + (idx_t)n * a]));
                ret = dd_add_eq(ret, dd_mul(v1, v2));
            }
        }
    } else {
        const int *ri = row + off[si], *cii = col + off[si];
        const dd *vi = val + off[si];
        const int ni = nnz[si];
        for (int q1 = 0; q1 < nj; q1++) {
            int alpha = rj[q1], beta = cj[q1];
            dd value1 = vj[q1];
            dd sum = mk(0.0);
            for (int q2 = 0; q2 < ni; q2++) {
                int gamma = ri[q2], delta = cii[q2];
                dd value2 = vi[q2];
                sum = dd_add_eq(sum, dd_mul(dd_mul(value2, Zinv[delta + (idx_t)n * beta]), X[alpha + (idx_t)n * gamma]));
                if (gamma != delta)
                    sum = dd_add_eq(sum, dd_mul(dd_mul(value2, Zinv[gamma + (idx_t)n * beta]), X[alpha + (idx_t)n * delta]));
            }
            ret = dd_add_eq(ret, dd_mul(value1, sum));
            if (alpha == beta) continue;
            sum = mk(0.0);
            for (int q2 = 0; q2 < ni; q2++) {
                int gamma = ri[q2], delta = cii[q2];
                dd value2 = vi[q2];
                sum = dd_add_eq(sum, dd_mul(dd_mul(value2, Zinv[delta + (idx_t)n * alpha]), X[beta + (idx_t)n * gamma]));
                if (gamma != delta)
                    sum = dd_add_eq(sum, dd_mul(dd_mul(value2, Zinv[gamma + (idx_t)n * alpha]), X[beta + (idx_t)n * delta]));
            }
            ret = dd_add_eq(ret, dd_mul(value1, sum));
        }
    }
    pv[p] = ret;
}
static constexpr int F2_FEW = 4;  // entries of A_j up to which the products are shared out
// F2 pairs with a sparse A_j (sdpa_newton.cpp:897-935), one workgroup per
// pair.  For each stored entry (a, b) of A_j the reference forms
// v = sum_t X(a,t) F(t,b) in t order (and the (b, a) sum for an
// off-diagonal entry) and adds v_j * v to the pair value.  The inner sums of
// different entries are independent: thread q forms those of entry q (its
// own chains in t order), and thread 0 then adds the terms in entry order.
__global__ __launch_bounds__(256) void k_pair_f2(int nlist, const int *list, int n, const int *psi, const int *psj,
                                                 const int *nnz, const idx_t *off, const int *wslot, const int *row,
                                                 const int *col, const dd *val, const dd *X, const dd *work1, dd *pv) {
    __shared__ dd va[256], vb[256];
    const int p = list[blockIdx.x], tid = threadIdx.x;
    const int si = psi[p], sj = psj[p];
    const int *rj = row + off[sj], *cj = col + off[sj];
    const dd *vj = val + off[sj];
    const int nj = nnz[sj];
    const dd *F = work1 + (idx_t)wslot[si] * n * n;
    dd ret = mk(0.0);
    if (nj <= F2_FEW) {
        // few entries (the usual case: an edge or a single coupling): every
        // thread forms products of the entries' sums, 256 t at a time, into
        // LDS, and thread d adds those of sum d in t order (the same
        // products and order as a thread forming its sum alone, without
        // one thread's serial loads along a row of X)
        __shared__ dd ps[2 * F2_FEW][256];
        __shared__ int sa[2 * F2_FEW], sb[2 * F2_FEW], snd;
        if (tid == 0) {
            int nd = 0;
            for (int q = 0; q < nj; q++) {
                const int a = rj[q], b = cj[q];
                sa[nd] = a;
                sb[nd++] = b;
                if (a != b) {
                    sa[nd] = b;
                    sb[nd++] = a;
                }
            }
            snd = nd;
        }
        __syncthreads();
        const int nd = snd;
        dd v = mk(0.0);  // sum tid (tid < nd)
        for (int t0 = 0; t0 < n; t0 += 256) {
            const int tc = n - t0 < 256 ? n - t0 : 256;
            for (int d = 0; d < nd; d++)
                if (tid < tc) {
                    const int t = t0 + tid;
                    ps[d][tid] = dd_mul(X[sa[d] + (idx_t)n * t], F[t + (idx_t)n * sb[d]]);
                }
            __syncthreads();
            if (tid < nd)
                for (int u = 0; u < tc; u++) v = dd_add_eq(v, ps[tid][u]);
            __syncthreads();
        }
        if (tid < nd) va[tid] = v;
        __syncthreads();
        if (tid == 0) {
            int d = 0;
            for (int q = 0; q < nj; q++) {
                const dd v1 = vj[q];
                ret = dd_add_eq(ret, dd_mul(v1, va[d++]));
                if (rj[q] != cj[q]) ret = dd_add_eq(ret, dd_mul(v1, va[d++]));
            }
            pv[p] = ret;
        }
        return;
    }
    for (int q0 = 0; q0 < nj; q0 += 256) {
        const int q = q0 + tid;
        if (q < nj) {
            const int a = rj[q], b = cj[q];
            dd v = mk(0.0);
#pragma unroll 8
            for (int t = 0; t < n; t++) v = dd_add_eq(v, dd_mul(X[a + (idx_t)n * t], F[t + (idx_t)n * b]));
            va[tid] = v;
            if (a != b) {
                v = mk(0.0);
#pragma unroll 8
                for (int t = 0; t < n; t++) v = dd_add_eq(v, dd_mul(X[b + (idx_t)n * t], F[t + (idx_t)n * a]));
                vb[tid] = v;
            }
        }
        __syncthreads();
        if (tid == 0) {
            const int cnt = nj - q0 < 256 ? nj - q0 : 256;
            for (int e = 0; e < cnt; e++) {
                const dd v1 = vj[q0 + e];
                ret = dd_add_eq(ret, dd_mul(v1, va[e]));
                if (rj[q0 + e] != cj[q0 + e]) ret = dd_add_eq(ret, dd_mul(v1, vb[e]));
            }
        }
        __syncthreads();
    }
    if (tid == 0) pv[p] = ret;
}
__global__ void k_scatter_pairs(int npairs, int m, const int *pi, const int *pj, const dd *pv, dd *B) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    int i = pi[p], j = pj[p];
    dd v = pv[p];
    if (i != j) {
        B[i + (idx_t)m * j] = dd_add_eq(B[i + (idx_t)m * j], v);
        B[j + (idx_t)m * i] = dd_add_eq(B[j + (idx_t)m * i], v);
    } else {
        B[i + (idx_t)m * i] = dd_add_eq(B[i + (idx_t)m * i], v);
    }
}
// B[e] = 0 + value of the pair that visits element e (the zeroed matrix and
// one "+=", as k_scatter_pairs gives), or 0; one SDP block only.
// B_ij = pair value pidx_ij (zero for -1), by 64 x 64 tiles: B_ij and B_ji
// read the same pair, and the pair values run contiguously along one of
// the two (the pair list is built row by row), so one half of B reads them
// strided.  Workgroups 16 q + x and 16 q + x + 8 (both on XCD x) take the
// tiles (a, b) and (b, a) of the u = 8 q + x-th tile pair a >= b, so a
// strided pair line is read by the workgroup beside the one that reads it
// contiguously, through the same L2.
__global__ __launch_bounds__(256) void k_gather_pairs(idx_t m, idx_t ntp, const int *pidx, const dd *pv, dd *B) {
    const idx_t L = blockIdx.x, u = (L / 16) * 8 + L % 8;
    const int h = (int)((L / 8) & 1);
    if (u >= ntp) return;
    idx_t a = (idx_t)((sqrt(8.0 * (double)u + 1.0) - 1.0) * 0.5);
    while ((a + 1) * (a + 2) / 2 <= u) a++;
    while (a * (a + 1) / 2 > u) a--;
    const idx_t b = u - a * (a + 1) / 2;
    if (h && a == b) return;
    const idx_t i = (h ? b : a) * 64 + (threadIdx.x & 63), j0 = (h ? a : b) * 64 + (threadIdx.x >> 6);
    if (i >= m) return;
#pragma unroll 4
    for (int r = 0; r < 16; r++) {
        const idx_t j = j0 + 4 * r;
        if (j >= m) break;
        const idx_t e = i + m * j;
        const int p = pidx[e];
        B[e] = p >= 0 ? dd_add_eq(mk(0.0), pv[p]) : mk(0.0);
    }
}
// Pointer-list GEMM support: fill pointer arrays for batched products.
__global__ void k_fill_ptrs(int cnt, const int *slots, const int *wslot, const int *type, const idx_t *off,
                            const dd *dense, dd *work1, dd *work2, idx_t nsq, int mode, const dd **A, const dd **B, dd **C,
                            const dd *X, const dd *Zinv) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cnt) return;
    int k = slots[q];
    if (mode == 0) {  // work1 = Adense * Zinv
        A[q] = dense + off[k];
        B[q] = Zinv;
        C[q] = work1 + (idx_t)wslot[k] * nsq;
    } else {  // work2 = X * work1
        A[q] = X;
        B[q] = work1 + (idx_t)wslot[k] * nsq;
        C[q] = work2 + (idx_t)wslot[k] * nsq;
    }
}

// LP part of the Schur matrix, after the SDP blocks: one thread per
// constraint pair i <= j adds x_l z_l^-1 a_i a_j (evaluated left to right,
// as compute_bMat_dense_LP does) for its LP entries in ascending l, to B_ij
// and, off the diagonal, to B_ji.
__global__ void k_schur_lp(int np, int m, const int *pi, const int *pj, const idx_t *ptr, const int *ll, const dd *ai,
                           const dd *aj, const dd *Xlp, const dd *Zilp, dd *B) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int i = pi[p], j = pj[p];
    dd bij = B[i + (idx_t)m * j], bji = B[j + (idx_t)m * i];
    for (idx_t q = ptr[p]; q < ptr[p + 1]; q++) {
        const int l = ll[q];
        dd v = dd_mul(dd_mul(dd_mul(Xlp[l], Zilp[l]), ai[q]), aj[q]);
        bij = dd_add_eq(bij, v);
        if (i != j) bji = dd_add_eq(bji, v);
    }
    B[i + (idx_t)m * j] = bij;
    if (i != j) B[j + (idx_t)m * i] = bji;
}

void schur(hipStream_t s, DeviceProblem &dp, const dd *X, const dd *invZ, dd *B, int T, bool lp_larger_first) {
    const int m = dp.m;
    // Newton::Make_bMat: bMat.setZero() (the gather below writes every element)
    if (!dp.d_pidx) sdk::fill(s, (idx_t)m * m, mk(0.0), B);
    for (size_t l = 0; l < dp.schur.size(); l++) {
        SchurBlock &S = dp.schur[l];
        const int n = S.n;
        const idx_t nsq = (idx_t)n * n;
        const dd *Xl = X + dp.boff[l];
        const dd *Zl = invZ + dp.boff[l];
        const dd **pA = (const dd **)dp.d_ptrs;
        const dd **pB = pA + dp.ptr_cap;
        dd **pC = (dd **)(pB + dp.ptr_cap);
        // work1 = A_i * Zinv
        if (!S.sp_slots.empty()) {
            dim3 g((unsigned)((nsq + 255) / 256), (unsigned)S.sp_slots.size());
            k_rowlist_mul<<<g, 256, 0, s>>>(n, (int)S.ci.size(), S.d_sp_slots, (int)S.sp_slots.size(), S.d_wslot,
                                            S.d_rl_ptr, S.d_rl_src, S.d_rl_val, Zl, dp.d_work1);
        }
        if (!S.de_slots.empty()) {
            int c = (int)S.de_slots.size();
            k_fill_ptrs<<<(c + 63) / 64, 64, 0, s>>>(c, S.d_de_slots, S.d_wslot, S.d_type, S.d_off, dp.d_dense, dp.d_work1,
                                                     dp.d_work2, nsq, 0, pA, pB, pC, Xl, Zl);
            sdk::gemm_ptrs(s, n, n, n, pA, n, pB, n, pC, n, c);
        }
        // work2 = X * work1 (F1, and F2 constraints paired with dense A_j)
        if (!S.g_slots.empty()) {
            int c = (int)S.g_slots.size();
            k_fill_ptrs<<<(c + 63) / 64, 64, 0, s>>>(c, S.d_g_slots, S.d_wslot, S.d_type, S.d_off, dp.d_dense, dp.d_work1,
                                                     dp.d_work2, nsq, 1, pA, pB, pC, Xl, Zl);
            sdk::gemm_ptrs(s, n, n, n, pA, n, pB, n, pC, n, c);
        }
        if (S.npairs > 0) {
            if (S.n_other > 0)
                k_pair_thread<<<(S.n_other + 127) / 128, 128, 0, s>>>(
                    S.n_other, S.d_list_other, n, S.d_pair_form, S.d_pair_slot_i, S.d_pair_slot_j, S.d_type, S.d_nnz,
                    S.d_off, S.d_wslot, dp.d_sp_row, dp.d_sp_col, dp.d_sp_val, Xl, Zl, dp.d_work1, dp.d_work2,
                    dp.d_pair_val);
            if (S.n_f2 > 0)
                k_pair_f2<<<S.n_f2, 256, 0, s>>>(S.n_f2, S.d_list_f2, n, S.d_pair_slot_i, S.d_pair_slot_j, S.d_nnz,
                                                S.d_off, S.d_wslot, dp.d_sp_row, dp.d_sp_col, dp.d_sp_val, Xl,
                                                dp.d_work1, dp.d_pair_val);
            if (S.ndense_pairs > 0) sdk::dot_batch(s, S.d_pair_tasks, S.ndense_pairs, T);
            if (dp.d_pidx) {
                const idx_t nt = ((idx_t)m + 63) / 64, ntp = nt * (nt + 1) / 2;
                k_gather_pairs<<<(unsigned)(16 * ((ntp + 7) / 8)), 256, 0, s>>>(m, ntp, dp.d_pidx, dp.d_pair_val, B);
            } else
                k_scatter_pairs<<<(S.npairs + 255) / 256, 256, 0, s>>>(S.npairs, m, S.d_pair_i, S.d_pair_j,
                                                                       dp.d_pair_val, B);
        }
    }
    // compute_bMat_dense_LP forms x z^-1 a_i a_j with i the earlier constraint
    // of a pair; compute_bMat_sparse_LP (sdpa_newton.cpp:1199-1219, pairs
    // kept when i >= j) with i the later one
    if (dp.n_lp_pairs > 0)
        k_schur_lp<<<(dp.n_lp_pairs + 255) / 256, 256, 0, s>>>(
            dp.n_lp_pairs, m, dp.d_lpp_i, dp.d_lpp_j, dp.d_lpp_ptr, dp.d_lpp_l,
            lp_larger_first ? dp.d_lpp_aj : dp.d_lpp_ai, lp_larger_first ? dp.d_lpp_ai : dp.d_lpp_aj, X + dp.nn,
            invZ + dp.nn, B);
}

}  // namespace sdp
