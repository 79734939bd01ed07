// ORACLE — test infrastructure only.
// CPU restatement of SDPA-DD's dense-Schur SDP Newton step and main loop.
// Reference files: sdpa_main.cpp, sdpa_newton.cpp, sdpa_parts.cpp,
// sdpa_dataset.cpp, sdpa_linear.cpp, sdpa_jordan.cpp, sdpa_io.cpp,
// sdpa_struct.cpp (cited per function).
#include "sdpa_oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <fstream>
#include <set>
#include <sstream>
#include <stdexcept>

namespace orc {

static const dd_real MONE = 1.0, MMONE = -1.0;

// ---------------------------------------------------------------- reader --
// sdpa_struct.cpp:458-498 (changeToDense, 20% threshold on stored entries).
void SparseMatrix::changeToDense() {
    if (type != SPARSE) return;
    if (NonZeroCount < (nRow * nCol) * 0.20) return;
    type = DENSE;
    size_t len = (size_t)nRow * nCol;
    de_ele.assign(len, dd_real(0.0));
    for (int idx = 0; idx < NonZeroCount; idx++) {
        int i = row_index[idx], j = column_index[idx];
        if (i == j)
            de_ele[i + (size_t)nCol * j] = sp_ele[idx];
        else
            de_ele[i + (size_t)nCol * j] = de_ele[j + (size_t)nCol * i] = sp_ele[idx];
    }
    NonZeroCount = NonZeroNumber = NonZeroEffect = (int)len;
    row_index.clear();
    column_index.clear();
    sp_ele.clear();
}

namespace {
// A cursor reproducing the fscanf patterns the reference reader uses.
struct Scan {
    const std::string &s;
    size_t p = 0;
    explicit Scan(const std::string &str) : s(str) {}
    static bool inset(char c) { return (c >= '0' && c <= '9') || c == '+' || c == '-'; }
    // "%*[^0-9+-]": at least one character outside the set.
    bool skip_sep() {
        size_t q = p;
        while (q < s.size() && !inset(s[q])) q++;
        if (q == p) return false;
        p = q;
        return p < s.size();
    }
    // "%d" (leading white space allowed).
    bool read_int(int &v) {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r' || s[p] == '\v' || s[p] == '\f')) p++;
        size_t q = p;
        if (q < s.size() && (s[q] == '+' || s[q] == '-')) q++;
        size_t d0 = q;
        while (q < s.size() && s[q] >= '0' && s[q] <= '9') q++;
        if (q == d0) return false;
        v = std::atoi(s.substr(p, q - p).c_str());
        p = q;
        return true;
    }
    // "%[^,} \t\n]".
    bool read_token(std::string &t) {
        size_t q = p;
        while (q < s.size() && s[q] != ',' && s[q] != '}' && s[q] != ' ' && s[q] != '\t' && s[q] != '\n') q++;
        if (q == p) return false;
        t = s.substr(p, q - p);
        p = q;
        return true;
    }
    bool sep_int(int &v) { return skip_sep() && read_int(v); }
    bool sep_token(std::string &t) { return skip_sep() && read_token(t); }
};

dd_real parse_value(const std::string &tok) {
    dd_real v;
    if (read_dd(tok.c_str(), v) != 0) v = dd_real(std::nan(""), std::nan(""));
    return v;
}
}  // namespace

bool read_dat_s(const std::string &path, InputData &data, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string buf = ss.str();
    Scan sc(buf);
    // Header: comment lines start with '*' or '"' (sdpa_io.cpp:124-135).
    int m = 0;
    for (;;) {
        if (sc.p >= buf.size()) {
            err = "no mDIM line";
            return false;
        }
        size_t e = buf.find('\n', sc.p);
        std::string line = buf.substr(sc.p, (e == std::string::npos ? buf.size() : e + 1) - sc.p);
        sc.p = (e == std::string::npos) ? buf.size() : e + 1;
        if (line[0] == '*' || line[0] == '"') continue;
        if (std::sscanf(line.c_str(), "%d", &m) != 1) {
            err = "bad mDIM";
            return false;
        }
        break;
    }
    int nBlock = 0;
    if (!sc.read_int(nBlock) || nBlock <= 0) {
        err = "bad nBLOCK";
        return false;
    }
    // Block types (sdpa_main.cpp:111-130): size >= 2 is an SDP block, a
    // negative size -k or the size 1 adds k (or 1) LP entries.
    std::vector<int> fbs(nBlock), btype(nBlock), bnum(nBlock), bs;
    int nLP = 0;
    for (int l = 0; l < nBlock; l++) {
        if (!sc.sep_int(fbs[l])) {
            err = "bad blockStruct";
            return false;
        }
        if (fbs[l] >= 2) {
            btype[l] = 1;
            bnum[l] = (int)bs.size();
            bs.push_back(fbs[l]);
        } else if (fbs[l] < 0 || fbs[l] == 1) {
            btype[l] = 3;
            bnum[l] = nLP;
            fbs[l] = fbs[l] < 0 ? -fbs[l] : 1;
            nLP += fbs[l];
        } else {
            err = "bad blockStruct";
            return false;
        }
    }
    const int nSDP = (int)bs.size();
    data.m = m;
    data.blockStruct = bs;
    data.LP_nBlock = nLP;
    data.b.assign(m, dd_real(0.0));
    for (int k = 0; k < m; k++) {
        std::string t;
        if (!sc.sep_token(t)) {
            err = "bad b vector";
            return false;
        }
        data.b[k] = parse_value(t);
    }
    // Entries "k l i j value" (sdpa_io.cpp:457-486 and 652-705).
    struct Ent {
        int k, l, i, j;
        dd_real v;
    };
    std::vector<Ent> ents;
    for (;;) {
        Ent e;
        std::string t;
        if (!sc.sep_int(e.k)) break;
        if (!sc.sep_int(e.l)) break;
        if (!sc.sep_int(e.i)) break;
        if (!sc.sep_int(e.j)) break;
        if (!sc.sep_token(t)) break;
        double chk;
        if (std::sscanf(t.c_str(), "%lf", &chk) <= 0) break;
        e.v = parse_value(t);
        if (e.k < 0 || e.k > m || e.l < 1 || e.l > nBlock) {
            err = "entry index out of range";
            return false;
        }
        ents.push_back(e);
    }
    // Blocks present per matrix, sorted (setBlockStruct), then elements in
    // file order (setElement).  LP entries must be diagonal; their indices
    // are sorted and deduplicated (sdpa_io.cpp:509-517, 607-617) and a
    // repeated entry keeps the last value (setElement_LP, sdpa_struct.cpp:1253-1265).
    std::vector<std::vector<int>> cnt(m + 1, std::vector<int>(nSDP, 0));
    std::vector<std::vector<int>> lpidx(m + 1);
    for (auto &e : ents) {
        if (btype[e.l - 1] == 1) {
            cnt[e.k][bnum[e.l - 1]]++;
        } else {
            if (e.i != e.j) {
                err = "LP entry off the diagonal";
                return false;
            }
            if (e.i < 1 || e.i > fbs[e.l - 1]) {
                err = "out of range in input data";
                return false;
            }
            lpidx[e.k].push_back(bnum[e.l - 1] + e.i - 1);
        }
    }
    auto init_space = [&](SparseLinearSpace &sp, int k) {
        std::vector<int> &li = lpidx[k];
        std::sort(li.begin(), li.end());
        li.erase(std::unique(li.begin(), li.end()), li.end());
        sp.LP_sp_index = li;
        sp.LP_sp_block.assign(li.size(), dd_real(0.0));
        for (int l = 0; l < nSDP; l++) {
            if (cnt[k][l] == 0) continue;
            sp.SDP_sp_index.push_back(l);
            SparseMatrix sm;
            sm.nRow = sm.nCol = bs[l];
            sm.type = SparseMatrix::SPARSE;
            sm.NonZeroNumber = cnt[k][l];
            sm.row_index.resize(cnt[k][l]);
            sm.column_index.resize(cnt[k][l]);
            sm.sp_ele.resize(cnt[k][l]);
            sp.SDP_sp_block.push_back(sm);
        }
    };
    init_space(data.C, 0);
    data.A.resize(m);
    for (int k = 0; k < m; k++) init_space(data.A[k], k + 1);
    for (auto &e : ents) {
        SparseLinearSpace &sp = (e.k == 0) ? data.C : data.A[e.k - 1];
        if (btype[e.l - 1] == 3) {
            int idx = bnum[e.l - 1] + e.i - 1, kk = 0;
            while (sp.LP_sp_index[kk] != idx) kk++;
            sp.LP_sp_block[kk] = (e.k == 0) ? -e.v : e.v;
            continue;
        }
        int l2 = bnum[e.l - 1], kk = 0;
        while (sp.SDP_sp_index[kk] != l2) kk++;
        SparseMatrix &sm = sp.SDP_sp_block[kk];
        if (e.i - 1 >= sm.nRow || e.j - 1 >= sm.nCol || e.i < 1 || e.j < 1) {
            err = "out of range in input data";
            return false;
        }
        int c = sm.NonZeroCount;
        sm.row_index[c] = e.i - 1;
        sm.column_index[c] = e.j - 1;
        sm.sp_ele[c] = (e.k == 0) ? -e.v : e.v;
        sm.NonZeroCount++;
        sm.NonZeroEffect += (e.i == e.j) ? 1 : 2;
    }
    // sdpa_main.cpp:197-200
    for (auto &b : data.C.SDP_sp_block) b.changeToDense();
    for (auto &a : data.A)
        for (auto &b : a.SDP_sp_block) b.changeToDense();
    data.initialize_index();
    return true;
}

// A value of an initial-point file: the exact pair HI+LO / HI-LO of two
// doubles, or a decimal read as QD reads it.
static dd_real parse_init_value(const std::string &tok) {
    const char *s = tok.c_str();
    char *e1 = nullptr;
    double hi = std::strtod(s, &e1);
    if (e1 != s && (*e1 == '+' || *e1 == '-')) {
        char *e2 = nullptr;
        double lo = std::strtod(e1, &e2);
        if (e2 != e1 && *e2 == '\0') return dd_real(hi, lo);
    }
    return parse_value(tok);
}

bool read_init(const std::string &path, const InputData &data, const Parameter &param, bool sparse, InitPoint &ip,
               std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string buf = ss.str();
    Scan sc(buf);
    const std::vector<int> &bs = data.blockStruct;
    const int nB = (int)bs.size(), nLP = data.LP_nBlock;
    std::vector<size_t> boff(nB + 1, 0);
    for (int l = 0; l < nB; l++) boff[l + 1] = boff[l] + (size_t)bs[l] * bs[l];
    const size_t nt = boff[nB] + nLP;
    // Solutions::initialize: X = Z = lambda I
    const dd_real lambda = param.lambdaStar;
    ip.x.assign(nt, dd_real(0.0));
    for (int l = 0; l < nB; l++)
        for (int i = 0; i < bs[l]; i++) ip.x[boff[l] + (size_t)i * bs[l] + i] = lambda;
    for (int i = 0; i < nLP; i++) ip.x[boff[nB] + i] = lambda;
    ip.z = ip.x;
    ip.y.assign(data.m, dd_real(0.0));
    for (int k = 0; k < data.m; k++) {  // yVec is read with the opposite sign
        std::string t;
        if (!sc.sep_token(t)) {
            err = "initial point: short y vector";
            return false;
        }
        ip.y[k] = -parse_init_value(t);
    }
    auto set_sdp = [&](std::vector<dd_real> &v, int l, int i, int j, const dd_real &val) {  // setElement_SDP
        v[boff[l] + i + (size_t)bs[l] * j] = val;
        v[boff[l] + j + (size_t)bs[l] * i] = val;
    };
    if (sparse) {
        for (;;) {
            int target, l, i, j;
            std::string t;
            if (!sc.sep_int(target) || !sc.sep_int(l) || !sc.sep_int(i) || !sc.sep_int(j) || !sc.sep_token(t)) break;
            double chk;
            if (std::sscanf(t.c_str(), "%lf", &chk) <= 0) break;
            dd_real val = parse_init_value(t);
            std::vector<dd_real> &v = (target == 1) ? ip.z : ip.x;
            if (l >= 1 && l <= nB) {
                if (i < 1 || j < 1 || i > bs[l - 1] || j > bs[l - 1]) {
                    err = "initial point: entry out of range";
                    return false;
                }
                set_sdp(v, l - 1, i - 1, j - 1, val);
            } else {
                if (i != j || i < 1 || i > nLP || l < 1) {
                    err = "initial point: bad LP entry";
                    return false;
                }
                v[boff[nB] + i - 1] = val;
            }
        }
    } else {
        for (int pass = 0; pass < 2; pass++) {  // Z, then X
            std::vector<dd_real> &v = pass ? ip.x : ip.z;
            for (int l = 0; l < nB; l++)
                for (int i = 0; i < bs[l]; i++)
                    for (int j = 0; j < bs[l]; j++) {
                        std::string t;
                        if (!sc.sep_token(t)) {
                            err = "initial point: short dense block";
                            return false;
                        }
                        dd_real val = parse_init_value(t);
                        if (i <= j && val != 0.0) set_sdp(v, l, i, j, val);
                    }
            for (int j = 0; j < nLP; j++) {
                std::string t;
                if (!sc.sep_token(t)) {
                    err = "initial point: short LP part";
                    return false;
                }
                dd_real val = parse_init_value(t);
                if (val != 0.0) v[boff[nB] + j] = val;
            }
        }
    }
    return true;
}

// sdpa_dataset.cpp:254-330
// and 395-465 (the LP entries' constraint lists)
void InputData::initialize_index() {
    int nB = (int)blockStruct.size();
    SDP_constraint.assign(nB, {});
    SDP_blockIndex.assign(nB, {});
    for (int i = 0; i < m; i++)
        for (size_t k = 0; k < A[i].SDP_sp_index.size(); k++) {
            int l = A[i].SDP_sp_index[k];
            SDP_constraint[l].push_back(i);
            SDP_blockIndex[l].push_back((int)k);
        }
    LP_constraint.assign(LP_nBlock, {});
    LP_blockIndex.assign(LP_nBlock, {});
    for (int i = 0; i < m; i++)
        for (size_t k = 0; k < A[i].LP_sp_index.size(); k++) {
            int l = A[i].LP_sp_index[k];
            LP_constraint[l].push_back(i);
            LP_blockIndex[l].push_back((int)k);
        }
}

// sdpa_parts.cpp:172-183
bool Parameter::readFile(const std::string &path) {
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) return false;
    int r = 0;
    r += std::fscanf(fp, "%d%*[^\n]", &maxIteration);
    r += std::fscanf(fp, "%lf%*[^\n]", &epsilonStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &lambdaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &omegaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &lowerBound);
    r += std::fscanf(fp, "%lf%*[^\n]", &upperBound);
    r += std::fscanf(fp, "%lf%*[^\n]", &betaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &betaBar);
    r += std::fscanf(fp, "%lf%*[^\n]", &gammaStar);
    r += std::fscanf(fp, "%lf%*[^\n]", &epsilonDash);
    std::fclose(fp);
    return r == 10;
}

// ------------------------------------------------------ linear algebra --
// sdpa_linear.cpp:202-298 (SparseMatrix . DenseMatrix).
static dd_real inner_sparse_dense(const SparseMatrix &a, const DenseMatrix &b) {
    if (a.type == SparseMatrix::DENSE) return Rdot((mint)a.nRow * a.nCol, a.de_ele.data(), 1, b.de_ele.data(), 1);
    const int nR = b.nRow;
    auto term = [&](int idx) {
        int i = a.row_index[idx], j = a.column_index[idx];
        const dd_real &v = a.sp_ele[idx];
        if (i == j) return v * b.de_ele[i + (size_t)nR * j];
        return v * (b.de_ele[i + (size_t)nR * j] + b.de_ele[j + (size_t)nR * i]);
    };
    dd_real ret = 0.0;
    int amari = a.NonZeroCount % 4, shou = a.NonZeroCount / 4;
    for (int idx = 0; idx < amari; idx++) ret += term(idx);
    for (int c = 0, idx = amari; c < shou; c++, idx += 4) {
        dd_real r1 = term(idx), r2 = term(idx + 1), r3 = term(idx + 2), r4 = term(idx + 3);
        ret += (r1 + r2 + r3 + r4);
    }
    return ret;
}
// sdpa_linear.cpp:185-200 and 1571-1614 (DenseLinearSpace . DenseLinearSpace).
static dd_real inner(const DenseLinearSpace &a, const DenseLinearSpace &b) {
    dd_real ret = 0.0;
    for (size_t l = 0; l < a.SDP_block.size(); l++) {
        const DenseMatrix &x = a.SDP_block[l];
        ret += Rdot((mint)x.nRow * x.nCol, x.de_ele.data(), 1, b.SDP_block[l].de_ele.data(), 1);
    }
    for (size_t l = 0; l < a.LP_block.size(); l++) ret += a.LP_block[l] * b.LP_block[l];
    return ret;
}
// sdpa_linear.cpp:1616-1650 (SparseLinearSpace . DenseLinearSpace).
static dd_real inner(const SparseLinearSpace &a, const DenseLinearSpace &b) {
    dd_real ret = 0.0;
    for (size_t l = 0; l < a.SDP_sp_index.size(); l++) ret += inner_sparse_dense(a.SDP_sp_block[l], b.SDP_block[a.SDP_sp_index[l]]);
    for (size_t l = 0; l < a.LP_sp_index.size(); l++) ret += a.LP_sp_block[l] * b.LP_block[a.LP_sp_index[l]];
    return ret;
}
static dd_real inner(const std::vector<dd_real> &a, const std::vector<dd_real> &b) {
    return Rdot((mint)a.size(), a.data(), 1, b.data(), 1);
}
// ret = scalar * a  (Lal::multiply, sdpa_linear.cpp:989-1008, 1652-1692)
static void scale(DenseLinearSpace &ret, const DenseLinearSpace &a, const dd_real &s) {
    for (size_t l = 0; l < a.SDP_block.size(); l++) {
        mint len = (mint)a.SDP_block[l].de_ele.size();
        if (&ret != &a) Rcopy(len, a.SDP_block[l].de_ele.data(), 1, ret.SDP_block[l].de_ele.data(), 1);
        Rscal(len, s, ret.SDP_block[l].de_ele.data(), 1);
    }
    for (size_t l = 0; l < a.LP_block.size(); l++) ret.LP_block[l] = a.LP_block[l] * s;
}
// ret = a + s*b  (Lal::plus dense, sdpa_linear.cpp:1108-1129, 1694-1737)
static void plus(DenseLinearSpace &ret, const DenseLinearSpace &a, const DenseLinearSpace &b, const dd_real &s) {
    for (size_t l = 0; l < a.SDP_block.size(); l++) {
        mint len = (mint)a.SDP_block[l].de_ele.size();
        if (&ret != &a) Rcopy(len, a.SDP_block[l].de_ele.data(), 1, ret.SDP_block[l].de_ele.data(), 1);
        Raxpy(len, s, b.SDP_block[l].de_ele.data(), 1, ret.SDP_block[l].de_ele.data(), 1);
    }
    for (size_t l = 0; l < a.LP_block.size(); l++) ret.LP_block[l] = a.LP_block[l] + b.LP_block[l] * s;
}
// ret += s*A for sparse A (Lal::plus(Dense, Dense, Sparse), sdpa_linear.cpp:1224-1317, 1777-1812);
// ret aliases a on every call site on the path.
static void plus_sparse(DenseLinearSpace &ret, const SparseLinearSpace &A, const dd_real &s) {
    for (size_t l = 0; l < A.SDP_sp_index.size(); l++) {
        DenseMatrix &R = ret.SDP_block[A.SDP_sp_index[l]];
        const SparseMatrix &B = A.SDP_sp_block[l];
        if (B.type == SparseMatrix::SPARSE) {
            for (int idx = 0; idx < B.NonZeroCount; idx++) {
                int i = B.row_index[idx], j = B.column_index[idx];
                dd_real value = B.sp_ele[idx] * s;
                if (i != j) {
                    R.de_ele[i + (size_t)R.nCol * j] += value;
                    R.de_ele[j + (size_t)R.nCol * i] += value;
                } else {
                    R.de_ele[i + (size_t)R.nCol * i] += value;
                }
            }
        } else {
            Raxpy((mint)R.de_ele.size(), s, B.de_ele.data(), 1, R.de_ele.data(), 1);
        }
    }
    for (size_t l = 0; l < A.LP_sp_index.size(); l++) {
        dd_real &r = ret.LP_block[A.LP_sp_index[l]];
        r = r + A.LP_sp_block[l] * s;
    }
}
static void gemm_nn(DenseMatrix &C, const DenseMatrix &A, const DenseMatrix &B) {
    Rgemm('N', 'N', C.nRow, C.nCol, A.nCol, MONE, A.de_ele.data(), A.nRow, B.de_ele.data(), B.nRow, 0.0, C.de_ele.data(),
          C.nRow);
}
// Jal::ns_jordan_triple_product (sdpa_jordan.cpp:335-339): ret = a*b*c.
static void triple(DenseLinearSpace &ret, const DenseLinearSpace &a, const DenseLinearSpace &b,
                   const DenseLinearSpace &c, DenseLinearSpace &work) {
    for (size_t l = 0; l < a.SDP_block.size(); l++) gemm_nn(work.SDP_block[l], a.SDP_block[l], b.SDP_block[l]);
    for (size_t l = 0; l < a.SDP_block.size(); l++) gemm_nn(ret.SDP_block[l], work.SDP_block[l], c.SDP_block[l]);
    for (size_t l = 0; l < a.LP_block.size(); l++) {
        work.LP_block[l] = a.LP_block[l] * b.LP_block[l];
        ret.LP_block[l] = work.LP_block[l] * c.LP_block[l];
    }
}
// sdpa_linear.cpp:832-909: ret = A * B for sparse A via row axpys.
static void mul_sparse_dense(DenseMatrix &ret, const SparseMatrix &A, const DenseMatrix &B) {
    for (auto &v : ret.de_ele) v = 0.0;
    if (A.type == SparseMatrix::DENSE) {
        Rgemm('N', 'N', ret.nRow, ret.nCol, A.nCol, MONE, A.de_ele.data(), A.nRow, B.de_ele.data(), B.nRow, 0.0,
              ret.de_ele.data(), ret.nRow);
        return;
    }
    for (int idx = 0; idx < A.NonZeroCount; idx++) {
        int i = A.row_index[idx], j = A.column_index[idx];
        dd_real value = A.sp_ele[idx];
        if (i != j) {
            Raxpy(B.nCol, value, B.de_ele.data() + j, B.nRow, ret.de_ele.data() + i, ret.nRow);
            Raxpy(B.nCol, value, B.de_ele.data() + i, B.nRow, ret.de_ele.data() + j, ret.nRow);
        } else {
            Raxpy(B.nCol, value, B.de_ele.data() + j, B.nRow, ret.de_ele.data() + j, ret.nRow);
        }
    }
}
// sdpa_linear.cpp:414-438
static void symmetrize(DenseMatrix &a) {
    const int n = a.nRow;
    for (int index = 0; index < n - 1; index++) {
        int index1 = index + index * n + 1, index2 = index + (index + 1) * n, length = n - 1 - index;
        Raxpy(length, MONE, &a.de_ele[index2], n, &a.de_ele[index1], 1);
        Rscal(length, dd_real(0.5), &a.de_ele[index1], 1);
        Rcopy(length, &a.de_ele[index1], 1, &a.de_ele[index2], n);
    }
}
// Lal::getCholesky + getInvLowTriangularMatrix (sdpa_linear.cpp:300-348, 397-412).
static bool inv_chol(DenseMatrix &invChol, const DenseMatrix &a, DenseMatrix &work) {
    const int n = a.nRow;
    Rcopy((mint)n * n, a.de_ele.data(), 1, work.de_ele.data(), 1);
    mint info = 0;
    Rpotrf_L(n, work.de_ele.data(), n, info);
    if (info != 0) return false;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < j; i++) work.de_ele[i + (size_t)n * j] = 0.0;
    invChol.setIdentity(MONE);
    Rtrsm_LLNN(n, n, MONE, work.de_ele.data(), n, invChol.de_ele.data(), n);
    return true;
}

// --------------------------------------------------------------- Lanczos --
// sdpa_linear.cpp:27-122: truncated Lanczos estimate of lambda_min(L M L^T).
dd_real lanczos_min_eigen(const DenseMatrix &L, const DenseMatrix &M) {
    const int nDim = M.nRow;
    dd_real alpha, beta, value;
    dd_real min = 1.0e+51, min_old = 1.0e+52, min_min = 1.0e+50;
    dd_real error = 1.0e+10;
    int k = 0, kk = 0;
    std::vector<dd_real> diagVec(nDim, min_min), diagVec2(nDim, dd_real(0.0)), q(nDim, dd_real(0.0)), r(nDim, MONE);
    std::vector<dd_real> qold(nDim), w(nDim), tmp(nDim), out(nDim), bv(nDim), workVec(std::max(1, 3 * nDim - 1));
    std::vector<dd_real> Q((size_t)nDim * nDim);
    beta = sqrt((dd_real)nDim);
    while (k < nDim && k < sqrt((dd_real)nDim) + 10 && beta > 1.0e-16 &&
           (abs(min - min_old) > (1.0e-5) * abs(min) + (1.0e-8) || abs(error * beta) > (1.0e-2) * abs(min) + (1.0e-4))) {
        qold = q;
        value = MONE / beta;
        Rcopy(nDim, r.data(), 1, q.data(), 1);
        Rscal(nDim, value, q.data(), 1);
        w = q;
        Rtrmv_LT(nDim, L.de_ele.data(), nDim, w.data());
        Rgemv_N(nDim, nDim, MONE, M.de_ele.data(), nDim, w.data(), 0.0, tmp.data());
        w = tmp;
        Rtrmv_LN(nDim, L.de_ele.data(), nDim, w.data());
        alpha = Rdot(nDim, q.data(), 1, w.data(), 1);
        diagVec[k] = alpha;
        Rcopy(nDim, w.data(), 1, r.data(), 1);
        Raxpy(nDim, -alpha, q.data(), 1, r.data(), 1);
        Raxpy(nDim, -beta, qold.data(), 1, r.data(), 1);
        if (kk >= sqrt((dd_real)k) || k == nDim - 1 || k > sqrt((dd_real)nDim + 9)) {
            kk = 0;
            out = diagVec;
            bv = diagVec2;
            out[nDim - 1] = diagVec[k];
            bv[nDim - 1] = 0.0;
            mint info;
            int kp1 = k + 1;
            Rsteqr('I', kp1, out.data(), bv.data(), Q.data(), nDim, workVec.data(), info);
            if (info < 0) {
                std::fprintf(stderr, "oracle Lanczos: bad argument\n");
                std::abort();
            } else if (info > 0) {
                break;
            }
            min_old = min;
            min = out[0];
            error = Q[k];
        }
        value = Rdot(nDim, r.data(), 1, r.data(), 1);
        beta = sqrt(value);
        diagVec2[k] = beta;
        ++k;
        ++kk;
    }
    return min - abs(error * beta);
}

// sdpa_jordan.cpp:62-66 and sdpa_linear.cpp:124-156: for a block of size
// <= 20, lambda_min(L M L^T) from all eigenvalues (Rsyev).
dd_real qr_min_eigen(const DenseMatrix &L, const DenseMatrix &M, mint &info) {
    const int n = M.nRow;
    std::vector<dd_real> t((size_t)n * n), a((size_t)n * n), w(n);
    // Lal::let(DLS2, '=', M, 'T', L) = M L^T, then Lal::let(DLS1, '=', L, '*', DLS2)
    Rgemm('N', 'T', n, n, n, MONE, M.de_ele.data(), n, L.de_ele.data(), n, 0.0, t.data(), n);
    Rgemm('N', 'N', n, n, n, MONE, L.de_ele.data(), n, t.data(), n, 0.0, a.data(), n);
    Rsyev_NL(n, a.data(), n, w.data(), info);
    return w[0];
}

// --------------------------------------------------------------- Schur --
enum Formula { F1 = 0, F2 = 1, F3 = 2 };

// sdpa_newton.cpp:694-819
std::vector<int> compute_formula(const InputData &data, dd_real Kappa) {
    const int m = data.m, nB = (int)data.blockStruct.size();
    std::vector<int> up(m * nB, 0), useFormula(m * nB, F1);
    for (int l = 0; l < nB; l++) {
        const auto &cons = data.SDP_constraint[l];
        const auto &bidx = data.SDP_blockIndex[l];
        for (size_t k1 = 0; k1 < cons.size(); k1++) {
            int i = cons[k1];
            int inz = data.A[i].SDP_sp_block[bidx[k1]].NonZeroEffect;
            int u = inz;
            for (size_t k2 = 0; k2 < cons.size(); k2++) {
                int j = cons[k2];
                int jnz = data.A[j].SDP_sp_block[bidx[k2]].NonZeroEffect;
                if (jnz < inz)
                    u += jnz;
                else if ((jnz == inz) && (j < i))
                    u += jnz;
            }
            up[i * nB + l] = u;
        }
    }
    for (int l = 0; l < nB; l++) {
        const auto &cons = data.SDP_constraint[l];
        const auto &bidx = data.SDP_blockIndex[l];
        for (size_t k = 0; k < cons.size(); k++) {
            int i = cons[k];
            const SparseMatrix &Ai = data.A[i].SDP_sp_block[bidx[k]];
            dd_real inz = Ai.NonZeroEffect;
            dd_real n = Ai.nRow;
            dd_real u = up[i * nB + l];
            dd_real f1 = Kappa * n * inz + n * n * n + Kappa * u;
            dd_real f2 = Kappa * n * inz + Kappa * (n + 1) * u;
            dd_real f3 = Kappa * (2 * Kappa * inz + 1) * u / Kappa;
            int f;
            if (Ai.type == SparseMatrix::DENSE)
                f = (f1 < f2) ? F1 : F2;
            else if (f1 < f2 && f1 < f3)
                f = F1;
            else if (f2 < f3)
                f = F2;
            else
                f = F3;
            useFormula[i * nB + l] = f;
        }
    }
    return useFormula;
}

// sdpa_newton.cpp:897-935
static dd_real calF2(const DenseMatrix &F, DenseMatrix &G, const DenseMatrix &X, const SparseMatrix &Aj, bool &hasG) {
    const int n = Aj.nRow;
    if (Aj.type == SparseMatrix::SPARSE) {
        dd_real ret = 0.0;
        for (int idx = 0; idx < Aj.NonZeroCount; idx++) {
            int a = Aj.row_index[idx], b = Aj.column_index[idx];
            dd_real v1 = Aj.sp_ele[idx];
            dd_real v2 = Rdot(n, X.de_ele.data() + a, n, F.de_ele.data() + (size_t)n * b, 1);
            ret += v1 * v2;
            if (a != b) {
                v2 = Rdot(n, X.de_ele.data() + b, n, F.de_ele.data() + (size_t)n * a, 1);
                ret += v1 * v2;
            }
        }
        return ret;
    }
    if (!hasG) {
        gemm_nn(G, X, F);
        hasG = true;
    }
    return inner_sparse_dense(Aj, G);
}

// sdpa_newton.cpp:937-977
static dd_real calF3(const DenseMatrix &X, const DenseMatrix &invZ, const SparseMatrix &Ai, const SparseMatrix &Aj) {
    dd_real ret = 0.0, sum;
    const int nz = invZ.nCol, nx = X.nCol;
    for (int i1 = 0; i1 < Aj.NonZeroCount; i1++) {
        int alpha = Aj.row_index[i1], beta = Aj.column_index[i1];
        dd_real value1 = Aj.sp_ele[i1];
        sum = 0.0;
        for (int i2 = 0; i2 < Ai.NonZeroCount; i2++) {
            int gamma = Ai.row_index[i2], delta = Ai.column_index[i2];
            dd_real value2 = Ai.sp_ele[i2];
            dd_real plu = value2 * invZ.de_ele[delta + (size_t)nz * beta] * X.de_ele[alpha + (size_t)nx * gamma];
            sum += plu;
            if (gamma != delta) {
                dd_real plu2 = value2 * invZ.de_ele[gamma + (size_t)nz * beta] * X.de_ele[alpha + (size_t)nx * delta];
                sum += plu2;
            }
        }
        ret += value1 * sum;
        if (alpha == beta) continue;
        sum = 0.0;
        for (int i2 = 0; i2 < Ai.NonZeroCount; i2++) {
            int gamma = Ai.row_index[i2], delta = Ai.column_index[i2];
            dd_real value2 = Ai.sp_ele[i2];
            dd_real plu = value2 * invZ.de_ele[delta + (size_t)nz * alpha] * X.de_ele[beta + (size_t)nx * gamma];
            sum += plu;
            if (gamma != delta) {
                dd_real plu2 = value2 * invZ.de_ele[gamma + (size_t)nz * alpha] * X.de_ele[beta + (size_t)nx * delta];
                sum += plu2;
            }
        }
        ret += value1 * sum;
    }
    return ret;
}

// sdpa_newton.cpp:979-1068 (B must be zero on entry: Make_bMat clears it).
void schur_dense(const InputData &data, const DenseLinearSpace &X, const DenseLinearSpace &invZ, std::vector<dd_real> &B,
                 const std::vector<int> &useFormula) {
    const int m = data.m, nB = (int)data.blockStruct.size();
    for (int l = 0; l < nB; l++) {
        const DenseMatrix &xMat = X.SDP_block[l], &invzMat = invZ.SDP_block[l];
        const auto &cons = data.SDP_constraint[l];
        const auto &bidx = data.SDP_blockIndex[l];
        // Rows are independent: row i writes only the pairs (i, j) it owns,
        // each exactly once per block, so a parallel loop over rows (each
        // thread with its own work matrices) gives the serial loop's values.
#pragma omp parallel
        {
        DenseMatrix work1, work2;
        work1.initialize(data.blockStruct[l]);
        work2.initialize(data.blockStruct[l]);
#pragma omp for schedule(dynamic, 1)
        for (size_t k1 = 0; k1 < cons.size(); k1++) {
            int i = cons[k1];
            const SparseMatrix &Ai = data.A[i].SDP_sp_block[bidx[k1]];
            int inz = Ai.NonZeroEffect;
            int formula = useFormula[i * nB + l];
            bool hasG = false;
            if (formula == F1) {
                mul_sparse_dense(work1, Ai, invzMat);
                gemm_nn(work2, xMat, work1);
            } else if (formula == F2) {
                mul_sparse_dense(work1, Ai, invzMat);
            }
            for (size_t k2 = 0; k2 < cons.size(); k2++) {
                int j = cons[k2];
                const SparseMatrix &Aj = data.A[j].SDP_sp_block[bidx[k2]];
                int jnz = Aj.NonZeroEffect;
                if ((inz < jnz) || ((inz == jnz) && (i < j))) continue;
                dd_real value;
                if (formula == F1)
                    value = inner_sparse_dense(Aj, work2);
                else if (formula == F2)
                    value = calF2(work1, work2, xMat, Aj, hasG);
                else
                    value = calF3(xMat, invzMat, Ai, Aj);
                if (i != j) {
                    B[i + (size_t)m * j] += value;
                    B[j + (size_t)m * i] += value;
                } else {
                    B[i + (size_t)m * i] += value;
                }
            }
        }
        }
    }
    // compute_bMat_dense_LP (sdpa_newton.cpp:1162-1197), after the SDP blocks
    for (int l = 0; l < (int)X.LP_block.size(); l++) {
        const dd_real &x = X.LP_block[l], &iz = invZ.LP_block[l];
        const auto &cons = data.LP_constraint[l];
        const auto &bidx = data.LP_blockIndex[l];
        for (size_t k1 = 0; k1 < cons.size(); k1++) {
            int i = cons[k1];
            const dd_real &Ai = data.A[i].LP_sp_block[bidx[k1]];
            for (size_t k2 = k1; k2 < cons.size(); k2++) {
                int j = cons[k2];
                const dd_real &Aj = data.A[j].LP_sp_block[bidx[k2]];
                dd_real value = x * iz * Ai * Aj;
                if (i != j) {
                    B[i + (size_t)m * j] += value;
                    B[j + (size_t)m * i] += value;
                } else {
                    B[i + (size_t)m * i] += value;
                }
            }
        }
    }
}

// ---------------------------------------------------------- sparse Schur --
// The chordal path of SDPA-C as SDPA-DD carries it: when the aggregate
// sparsity of B is low (Chordal::ordering_bMat, sdpa_chordal.cpp:348-440),
// B is held as the upper-triangular pattern of its factor in a
// fill-reducing order (Newton::initialize_sparse_bMat, sdpa_newton.cpp:
// 225-336), assembled pair by pair into that pattern
// (make_aggrigateIndex_SDP/_LP, sdpa_newton.cpp:398-455, 597-665;
// compute_bMat_sparse_SDP/_LP, 1070-1142, 1199-1219), factored by
// Lal::getCholesky (sdpa_linear.cpp:351-395) and solved by
// Lal::solveSystems (sparse, 654-790) between permuteVec and
// reverse_permuteVec (sdpa_newton.cpp:1257-1275).  The reference orders with
// Spooles 2.2's minimum degree, which is absent here; the ordering is the
// exact minimum-degree elimination of min_degree_order.
int g_schur_mode = 0;

// Chordal::makeGraph (sdpa_chordal.cpp:141-238): vertex i is adjacent to
// every constraint that shares an SDP block or an LP entry with it, itself
// included; lists ascending without repeats.
static std::vector<std::vector<int>> aggregate_graph(const InputData &d) {
    std::vector<std::vector<int>> adj(d.m);
    auto add = [&](const std::vector<int> &cons) {
        for (int i : cons) adj[i].insert(adj[i].end(), cons.begin(), cons.end());
    };
    for (const auto &c : d.SDP_constraint) add(c);
    for (const auto &c : d.LP_constraint) add(c);
    for (auto &a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    return adj;
}

// Exact minimum degree: repeatedly eliminate the vertex with the fewest
// remaining neighbours (ties: the smaller index); its remaining neighbours
// become a clique.  Returns newToOld; rows[p] holds the new indices of the
// neighbours of the p-th eliminated vertex at its elimination, ascending
// (the strict upper part of row p of the factor).
std::vector<int> min_degree_order(int m, const std::vector<std::vector<int>> &adj, std::vector<std::vector<int>> &rows) {
    std::vector<std::set<int>> g(m);
    for (int i = 0; i < m; i++)
        for (int j : adj[i])
            if (j != i) g[i].insert(j);
    std::vector<int> order, newIdx(m, -1);
    std::vector<bool> gone(m, false);
    std::vector<std::vector<int>> nb_old(m);
    for (int step = 0; step < m; step++) {
        int v = -1;
        size_t best = 0;
        for (int i = 0; i < m; i++)
            if (!gone[i] && (v < 0 || g[i].size() < best)) {
                v = i;
                best = g[i].size();
            }
        order.push_back(v);
        newIdx[v] = step;
        gone[v] = true;
        nb_old[step].assign(g[v].begin(), g[v].end());
        for (int u : g[v]) {
            g[u].erase(v);
            for (int w : g[v])
                if (w != u) g[u].insert(w);
        }
        g[v].clear();
    }
    rows.assign(m, {});
    for (int p = 0; p < m; p++) {
        for (int u : nb_old[p]) rows[p].push_back(newIdx[u]);
        std::sort(rows[p].begin(), rows[p].end());
    }
    return order;
}

// Decision, storage and pair locations.  mode 0: the reference's thresholds
// (m > 100, more than 5 blocks, every block's constraint count <= m
// sqrt(0.25), aggregate pattern <= 0.25 m^2, factor pattern 2 nnz - m <=
// 0.4 m^2); mode 2: the thresholds of its debugging switch (0, 0, 2.0,
// 2.0, sdpa_chordal.cpp:53-58), which take the sparse path whenever the
// pattern allows; mode 1: dense.
bool sparse_schur_setup(const InputData &d, SparseSchur &S, int mode) {
    S = SparseSchur();
    if (mode == 1) return false;
    const int m = d.m;
    const int nBlock2 = (int)d.blockStruct.size() + d.LP_nBlock;
    const int m_thr = mode == 2 ? 0 : 100, b_thr = mode == 2 ? 0 : 5;
    const double agg = mode == 2 ? 2.0 : 0.25, ext = mode == 2 ? 2.0 : 0.4;
    if (m <= m_thr || nBlock2 <= b_thr) return false;
    for (const auto &c : d.SDP_constraint)
        if ((int)c.size() > m * std::sqrt(agg)) return false;
    for (const auto &c : d.LP_constraint)
        if ((int)c.size() > m * std::sqrt(agg)) return false;
    const auto adj = aggregate_graph(d);
    long long tsize = 0;
    for (const auto &a : adj) tsize += (long long)a.size();
    if (tsize > agg * m * m) return false;
    std::vector<std::vector<int>> rows;
    std::vector<int> order = min_degree_order(m, adj, rows);
    long long nonzeros = 0;
    for (const auto &r : rows) nonzeros += 1 + (long long)r.size();
    if (nonzeros * 2 - m > ext * m * m) return false;
    S.on = true;
    S.m = m;
    S.ordering = order;
    S.reverse_ordering.assign(m, 0);
    for (int i = 0; i < m; i++) S.reverse_ordering[order[i]] = i;
    S.diagonalIndex.assign(m + 1, 0);
    for (int p = 0; p < m; p++) S.diagonalIndex[p + 1] = S.diagonalIndex[p] + 1 + (int)rows[p].size();
    const int nnz = S.diagonalIndex[m];
    S.row_index.resize(nnz);
    S.column_index.resize(nnz);
    S.sp_ele.assign(nnz, dd_real(0.0));
    for (int p = 0; p < m; p++) {
        int k = S.diagonalIndex[p];
        S.row_index[k] = p;
        S.column_index[k] = p;
        for (int c : rows[p]) {
            k++;
            S.row_index[k] = p;
            S.column_index[k] = c;
        }
    }
    auto locate = [&](int i, int j) {
        int ii = S.reverse_ordering[i], jj = S.reverse_ordering[j];
        if (ii > jj) std::swap(ii, jj);
        const int *b = S.column_index.data() + S.diagonalIndex[ii], *e = S.column_index.data() + S.diagonalIndex[ii + 1];
        const int *f = std::lower_bound(b, e, jj);
        if (f == e || *f != jj) throw std::runtime_error("sparse Schur: pair outside the factor pattern");
        return (int)(f - S.column_index.data());
    };
    const int nB = (int)d.blockStruct.size();
    S.sdp.resize(nB);
    for (int l = 0; l < nB; l++) {
        const auto &cons = d.SDP_constraint[l];
        const auto &bidx = d.SDP_blockIndex[l];
        for (size_t k1 = 0; k1 < cons.size(); k1++) {
            const int i = cons[k1], inz = d.A[i].SDP_sp_block[bidx[k1]].NonZeroEffect;
            for (size_t k2 = 0; k2 < cons.size(); k2++) {
                const int j = cons[k2], jnz = d.A[j].SDP_sp_block[bidx[k2]].NonZeroEffect;
                if ((inz < jnz) || ((inz == jnz) && (i < j))) continue;
                S.sdp[l].push_back({i, j, (int)k1, (int)k2, locate(i, j)});
            }
        }
    }
    S.lp.resize(d.LP_nBlock);
    for (int l = 0; l < d.LP_nBlock; l++) {
        const auto &cons = d.LP_constraint[l];
        for (size_t k1 = 0; k1 < cons.size(); k1++)
            for (size_t k2 = 0; k2 < cons.size(); k2++) {
                const int i = cons[k1], j = cons[k2];
                if (i < j) continue;
                S.lp[l].push_back({i, j, (int)k1, (int)k2, locate(i, j)});
            }
    }
    return true;
}

// compute_bMat_sparse_SDP then compute_bMat_sparse_LP (the pattern starts at zero).
void schur_sparse(const InputData &data, const DenseLinearSpace &X, const DenseLinearSpace &invZ, SparseSchur &S,
                  const std::vector<int> &useFormula) {
    for (auto &v : S.sp_ele) v = 0.0;
    const int nB = (int)data.blockStruct.size();
    for (int l = 0; l < nB; l++) {
        const DenseMatrix &xMat = X.SDP_block[l], &invzMat = invZ.SDP_block[l];
        const auto &bidx = data.SDP_blockIndex[l];
        DenseMatrix work1, work2;
        work1.initialize(data.blockStruct[l]);
        work2.initialize(data.blockStruct[l]);
        int previous_i = -1, formula = F3;
        bool hasG = false;
        for (const auto &pr : S.sdp[l]) {
            const SparseMatrix &Ai = data.A[pr.i].SDP_sp_block[bidx[pr.k1]];
            if (pr.i != previous_i) {
                formula = useFormula[pr.i * nB + l];
                hasG = false;
                if (formula == F1) {
                    mul_sparse_dense(work1, Ai, invzMat);
                    gemm_nn(work2, xMat, work1);
                } else if (formula == F2) {
                    mul_sparse_dense(work1, Ai, invzMat);
                }
            }
            const SparseMatrix &Aj = data.A[pr.j].SDP_sp_block[bidx[pr.k2]];
            dd_real value;
            if (formula == F1)
                value = inner_sparse_dense(Aj, work2);
            else if (formula == F2)
                value = calF2(work1, work2, xMat, Aj, hasG);
            else
                value = calF3(xMat, invzMat, Ai, Aj);
            S.sp_ele[pr.loc] += value;
            previous_i = pr.i;
        }
    }
    for (int l = 0; l < (int)X.LP_block.size(); l++) {
        const dd_real &x = X.LP_block[l], &iz = invZ.LP_block[l];
        const auto &bidx = data.LP_blockIndex[l];
        for (const auto &pr : S.lp[l]) {
            const dd_real &Ai = data.A[pr.i].LP_sp_block[bidx[pr.k1]];
            const dd_real &Aj = data.A[pr.j].LP_sp_block[bidx[pr.k2]];
            dd_real value = x * iz * Ai * Aj;
            S.sp_ele[pr.loc] += value;
        }
    }
}

// Lal::getCholesky(SparseMatrix&, int*) (sdpa_linear.cpp:351-395): row by
// row, the diagonal replaced by 1/sqrt (0 for a negative pivot, and the run
// goes on), the row scaled by it, then the row's outer product subtracted
// from the rows of its columns.
void chol_sparse(SparseSchur &S) {
    auto &a = S.sp_ele;
    const auto &col = S.column_index;
    for (int i = 0; i < S.m; ++i) {
        const int indexA1 = S.diagonalIndex[i], indexA2 = S.diagonalIndex[i + 1];
        if (a[indexA1] < 0.0)
            a[indexA1] = 0.0;
        else
            a[indexA1] = 1.0 / sqrt(a[indexA1]);
        for (int k1 = indexA1 + 1; k1 < indexA2; ++k1) a[k1] *= a[indexA1];
        for (int k1 = indexA1 + 1; k1 < indexA2; ++k1) {
            const dd_real tmp = a[k1];
            int k3 = S.diagonalIndex[col[k1]];
            const int indexB2 = S.diagonalIndex[col[k1] + 1];
            for (int k2 = k1; k2 < indexA2; ++k2) {
                const dd_real tmp2 = a[k2];
                const int tmp3 = col[k2];
                for (; k3 < indexB2; ++k3)
                    if (col[k3] == tmp3) {
                        a[k3] -= tmp * tmp2;
                        k3++;
                        break;
                    }
            }
        }
    }
}

// permuteVec, Lal::solveSystems (sparse) and reverse_permuteVec.
void solve_sparse(const SparseSchur &S, const std::vector<dd_real> &g, std::vector<dd_real> &dy) {
    const int m = S.m, nnz = S.diagonalIndex[m];
    std::vector<dd_real> x(m);
    for (int k = 0; k < m; k++) x[k] = g[S.ordering[k]];
    for (int index = 0; index < nnz; ++index) {
        const int i = S.row_index[index], j = S.column_index[index];
        const dd_real value = S.sp_ele[index];
        if (i == j)
            x[i] *= value;
        else
            x[j] -= value * x[i];
    }
    for (int index = nnz - 1; index >= 0; --index) {
        const int i = S.row_index[index], j = S.column_index[index];
        const dd_real value = S.sp_ele[index];
        if (i == j)
            x[i] *= value;
        else
            x[i] -= value * x[j];
    }
    dy.assign(m, dd_real(0.0));
    for (int k = 0; k < m; k++) dy[S.ordering[k]] = x[k];
}

// ---------------------------------------------------------------- solver --
namespace {
struct Solutions {
    int nDim = 0, mDim = 0;
    DenseLinearSpace xMat, zMat, invCholeskyX, invCholeskyZ, invzMat;
    std::vector<dd_real> yVec;
};
struct Residuals {
    std::vector<dd_real> primalVec;
    DenseLinearSpace dualMat;
    dd_real normPrimalVec, normDualMat;
};
struct Newton {
    std::vector<dd_real> gVec, DyVec, bMat;
    DenseLinearSpace DxMat, DzMat, r_zinvMat, x_rd_zinvMat;
    std::vector<int> useFormula;
};
struct Work {
    DenseLinearSpace DLS1, DLS2;
};

// sdpa_dataset.cpp:477-493
void multi_inner(const InputData &d, const DenseLinearSpace &x, std::vector<dd_real> &ret) {
    for (int i = 0; i < d.m; i++) ret[i] = inner(d.A[i], x);
}
void multi_plus(const InputData &d, const std::vector<dd_real> &y, DenseLinearSpace &ret) {
    ret.setZero();
    for (int i = 0; i < d.m; i++) plus_sparse(ret, d.A[i], y[i]);
}
// sdpa_dataset.cpp:583-611
dd_real max_norm(const std::vector<dd_real> &v) {
    dd_real ret = 0.0;
    for (auto &x : v) {
        dd_real t = abs(x);
        if (t > ret) ret = t;
    }
    return ret;
}
dd_real max_norm(const DenseLinearSpace &d) {
    dd_real ret = 0.0;
    for (auto &b : d.SDP_block)
        for (auto &x : b.de_ele) {
            dd_real t = abs(x);
            if (t > ret) ret = t;
        }
    for (auto &x : d.LP_block) {
        dd_real t = abs(x);
        if (t > ret) ret = t;
    }
    return ret;
}
// sdpa_dataset.cpp:625-645
void residuals_compute(const InputData &d, const Solutions &pt, Residuals &res) {
    multi_inner(d, pt.xMat, res.primalVec);
    Rscal(d.m, MMONE, res.primalVec.data(), 1);
    Raxpy(d.m, MONE, d.b.data(), 1, res.primalVec.data(), 1);
    multi_plus(d, pt.yVec, res.dualMat);
    scale(res.dualMat, res.dualMat, MMONE);
    plus_sparse(res.dualMat, d.C, MONE);
    plus(res.dualMat, res.dualMat, pt.zMat, MMONE);
    res.normPrimalVec = max_norm(res.primalVec);
    res.normDualMat = max_norm(res.dualMat);
}
// sdpa_dataset.cpp:96-118
// Jal::getInvChol / getInvCholAndInv (sdpa_jordan.cpp:123-170): an LP entry
// fails only when negative; its inverse factor is 1/sqrt(x), its inverse 1/z.
static bool inv_chol_space(DenseLinearSpace &invChol, const DenseLinearSpace &a, DenseLinearSpace &work) {
    for (size_t l = 0; l < a.SDP_block.size(); l++)
        if (!inv_chol(invChol.SDP_block[l], a.SDP_block[l], work.SDP_block[l])) return false;
    for (size_t l = 0; l < a.LP_block.size(); l++) {
        if (a.LP_block[l] < 0.0) return false;
        invChol.LP_block[l] = 1.0 / sqrt(a.LP_block[l]);
    }
    return true;
}
bool compute_inverse(Solutions &pt, Work &w) {
    bool ok = inv_chol_space(pt.invCholeskyX, pt.xMat, w.DLS1);
    bool okz = inv_chol_space(pt.invCholeskyZ, pt.zMat, w.DLS2);
    if (okz) {
        for (size_t l = 0; l < pt.zMat.SDP_block.size(); l++) {
            DenseMatrix &inv = pt.invzMat.SDP_block[l];
            const DenseMatrix &lc = pt.invCholeskyZ.SDP_block[l];
            inv.de_ele = lc.de_ele;
            Rtrmm_LLTN(lc.nRow, lc.nCol, MONE, lc.de_ele.data(), lc.nRow, inv.de_ele.data(), inv.nRow);
        }
        for (size_t l = 0; l < pt.zMat.LP_block.size(); l++) pt.invzMat.LP_block[l] = 1.0 / pt.zMat.LP_block[l];
    }
    return ok && okz;
}
// sdpa_jordan.cpp:55-88: Lanczos for SDP blocks of size > 20, Rsyev for
// the others, x l l for LP entries.  A failed Rsyev ends the reference's
// process (sdpa_linear.cpp:139-145); ok reports it here.
dd_real min_eigen(const DenseLinearSpace &L, const DenseLinearSpace &M, bool &ok) {
    dd_real min = 1.0E50;
    for (size_t l = 0; l < M.SDP_block.size(); l++) {
        dd_real v;
        if (M.SDP_block[l].nRow > 20) {
            v = lanczos_min_eigen(L.SDP_block[l], M.SDP_block[l]);
        } else {
            mint info = 0;
            v = qr_min_eigen(L.SDP_block[l], M.SDP_block[l], info);
            if (info != 0) ok = false;
        }
        if (v < min) min = v;
    }
    for (size_t l = 0; l < M.LP_block.size(); l++) {
        dd_real v = M.LP_block[l] * L.LP_block[l] * L.LP_block[l];
        if (v < min) min = v;
    }
    return min;
}
}  // namespace

bool solve(const InputData &data, const Parameter &param, SolveResult &res, const InitPoint *init) {
    const auto t_start = std::chrono::steady_clock::now();
    const int m = data.m;
    const std::vector<int> &bs = data.blockStruct;
    // Newton and work space (sdpa_newton.cpp:191-223, sdpa_dataset.cpp:668-702).
    Newton nw;
    nw.gVec.assign(m, dd_real(0.0));
    nw.DyVec.assign(m, dd_real(0.0));
    nw.bMat.assign((size_t)m * m, dd_real(0.0));
    nw.DxMat.initialize(bs, data.LP_nBlock);
    nw.DzMat.initialize(bs, data.LP_nBlock);
    nw.r_zinvMat.initialize(bs, data.LP_nBlock);
    nw.x_rd_zinvMat.initialize(bs, data.LP_nBlock);
    nw.useFormula = compute_formula(data, dd_real(1.2));
    // Chordal::ordering_bMat + Newton::initialize_bMat (sdpa_main.cpp:217-221)
    SparseSchur sps;
    const bool sparse = sparse_schur_setup(data, sps, g_schur_mode);
    if (sparse) {
        nw.bMat.clear();
        res.schur_sparse = true;
        res.schur_ordering = sps.ordering;
        res.schur_nnz = sps.diagonalIndex[m];
    }
    Work work;
    work.DLS1.initialize(bs, data.LP_nBlock);
    work.DLS2.initialize(bs, data.LP_nBlock);
    // Solutions::initialize (sdpa_dataset.cpp:35-60).
    Solutions pt;
    pt.mDim = m;
    for (int n : bs) pt.nDim += n;
    pt.nDim += data.LP_nBlock;
    dd_real lambda = param.lambdaStar;
    pt.xMat.initialize(bs, data.LP_nBlock);
    pt.zMat.initialize(bs, data.LP_nBlock);
    pt.invCholeskyX.initialize(bs, data.LP_nBlock);
    pt.invCholeskyZ.initialize(bs, data.LP_nBlock);
    pt.invzMat.initialize(bs, data.LP_nBlock);
    pt.xMat.setIdentity(lambda);
    pt.zMat.setIdentity(lambda);
    pt.invCholeskyX.setIdentity(1.0 / sqrt(lambda));
    pt.invCholeskyZ.setIdentity(1.0 / sqrt(lambda));
    pt.invzMat.setIdentity(1.0 / lambda);
    pt.yVec.assign(m, dd_real(0.0));
    const int nDim = pt.nDim;
    // sdpa_main.cpp:237-251: the initial point, its inverse factors (the
    // reference does not check computeInverse's result here) and copies
    DenseLinearSpace initX, initZ;
    auto load = [&](DenseLinearSpace &d, const std::vector<dd_real> &v) {
        size_t o = 0;
        for (auto &b : d.SDP_block)
            for (auto &e : b.de_ele) e = v[o++];
        for (auto &e : d.LP_block) e = v[o++];
    };
    if (init) {
        load(pt.xMat, init->x);
        load(pt.zMat, init->z);
        pt.yVec = init->y;
        (void)compute_inverse(pt, work);
        initX = pt.xMat;
        initZ = pt.zMat;
    }

    Residuals initRes;
    initRes.primalVec.assign(m, dd_real(0.0));
    initRes.dualMat.initialize(bs, data.LP_nBlock);
    residuals_compute(data, pt, initRes);
    Residuals cur = initRes;

    // sdpa_parts.cpp state
    dd_real alphaP = 0.0, alphaD = 0.0;                     // StepLength
    dd_real betaV = param.betaStar;                           // DirectionParameter
    bool switchOn = true;                                     // Switch
    dd_real muInitial = dd_real(param.lambdaStar) * dd_real(param.lambdaStar), muCurrent = muInitial;
    if (init) {  // AverageComplementarity::initialize(Solutions&) (sdpa_parts.cpp:583-588)
        muInitial = inner(pt.xMat, pt.zMat);
        muInitial /= nDim;
        muCurrent = muInitial;
    }
    dd_real thetaP, thetaD;                                   // RatioInitResCurrentRes
    {
        dd_real accuracy = param.epsilonDash;
        thetaP = (initRes.normPrimalVec < accuracy) ? dd_real(0.0) : dd_real(1.0);
        thetaD = (initRes.normDualMat < accuracy) ? dd_real(0.0) : dd_real(1.0);
    }
    // SolveInfo::initialize (sdpa_parts.cpp:688-699)
    dd_real rho = 1.0, etaPrimal, etaDual, objValPrimal, objValDual;
    {
        dd_real omegaStar = param.omegaStar;
        etaPrimal = omegaStar * nDim * muInitial;
        etaDual = omegaStar * nDim * muInitial;
        objValPrimal = inner(data.C, pt.xMat);
        objValDual = inner(data.b, pt.yVec);
    }
    int phase = noINFO;
    // Phase::updateCheck (sdpa_parts.cpp:875-954)
    auto updateCheck = [&](const Residuals &r) -> bool {
        const dd_real NONZERO = 1.0e-6;
        dd_real accuracy = param.epsilonDash;
        phase = noINFO;
        if (r.normPrimalVec <= accuracy) phase = (r.normDualMat <= accuracy) ? pdFEAS : pFEAS;
        if (phase == noINFO && r.normDualMat <= accuracy) phase = dFEAS;
        if (phase == pdFEAS) {
            dd_real mean = (abs(objValPrimal) + abs(objValDual)) / 2.0;
            dd_real PDgap = abs(objValPrimal - objValDual);
            dd_real dominator = (mean < 1.0) ? dd_real(1.0) : mean;
            if (PDgap / dominator <= param.epsilonStar) {
                phase = pdOPT;
                return false;
            }
        }
        if (phase == noINFO && rho > 1.0 + NONZERO) {
            phase = pdINF;
            return false;
        }
        if (phase == pFEAS) {
            if (objValPrimal <= -param.upperBound) {
                phase = pUNBD;
                return false;
            }
            if (rho > 1.0 + NONZERO) {
                phase = pFEAS_dINF;
                return false;
            }
        }
        if (phase == dFEAS) {
            if (objValDual >= -param.lowerBound) {
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
    updateCheck(initRes);

    // Newton::Mehrotra (sdpa_newton.cpp:1345-1384)
    auto mehrotra = [&](bool predictor) -> bool {
        // compute_rMat (821-837)
        dd_real target = betaV * muCurrent;
        scale(nw.r_zinvMat, pt.invzMat, target);
        plus(nw.r_zinvMat, nw.r_zinvMat, pt.xMat, MMONE);
        if (!predictor) {
            triple(work.DLS1, nw.DxMat, nw.DzMat, pt.invzMat, work.DLS2);
            plus(nw.r_zinvMat, nw.r_zinvMat, work.DLS1, MMONE);
        }
        // Make_gVec (839-893)
        if (phase == pFEAS || phase == noINFO) {
            if (!predictor) {
                plus(work.DLS1, nw.r_zinvMat, nw.x_rd_zinvMat, MMONE);
            } else {
                triple(nw.x_rd_zinvMat, pt.xMat, cur.dualMat, pt.invzMat, work.DLS2);
                plus(work.DLS1, nw.r_zinvMat, nw.x_rd_zinvMat, MMONE);
            }
        } else {
            work.DLS1 = nw.r_zinvMat;
        }
        multi_inner(data, work.DLS1, nw.gVec);
        Rscal(m, MMONE, nw.gVec.data(), 1);
        Raxpy(m, MONE, cur.primalVec.data(), 1, nw.gVec.data(), 1);
        if (predictor) {
            // Make_bMat (1223-1244) + compute_DyVec's Cholesky (1280-1296)
            if (sparse) {
                schur_sparse(data, pt.xMat, pt.invzMat, sps, nw.useFormula);
                chol_sparse(sps);
            } else {
                for (auto &v : nw.bMat) v = 0.0;
                schur_dense(data, pt.xMat, pt.invzMat, nw.bMat, nw.useFormula);
                mint info = 0;
                Rpotrf_L(m, nw.bMat.data(), m, info);
                if (info != 0) return false;
            }
        }
        if (sparse) {
            // compute_DyVec (sdpa_newton.cpp:1298-1302)
            solve_sparse(sps, nw.gVec, nw.DyVec);
        } else {
            // solveSystems (sdpa_linear.cpp:639-651)
            nw.DyVec = nw.gVec;
            Rtrsv_LN(m, nw.bMat.data(), m, nw.DyVec.data());
            Rtrsv_LT(m, nw.bMat.data(), m, nw.DyVec.data());
        }
        // compute_DzMat (1318-1327)
        multi_plus(data, nw.DyVec, nw.DzMat);
        scale(nw.DzMat, nw.DzMat, MMONE);
        if (phase == pFEAS || phase == noINFO) plus(nw.DzMat, nw.DzMat, cur.dualMat, MONE);
        // compute_DxMat (1329-1343)
        triple(work.DLS1, pt.xMat, nw.DzMat, pt.invzMat, work.DLS2);
        plus(nw.DxMat, nw.r_zinvMat, work.DLS1, MMONE);
        for (auto &blk : nw.DxMat.SDP_block) symmetrize(blk);
        return true;
    };

    int pIteration = 0;
    while (updateCheck(cur) && pIteration < param.maxIteration) {
        // Switch::MehrotraPredictor, DirectionParameter::MehrotraPredictor (sdpa_parts.cpp:486-555)
        switchOn = (phase == noINFO || phase == pFEAS || phase == dFEAS);
        if (phase == pdFEAS) {
            betaV = 0.0;
        } else {
            betaV = param.betaBar;
            if (!switchOn) betaV = dd_real(2.0);
        }
        if (!mehrotra(true)) {
            res.message = "Schur Cholesky failed";
            break;
        }
        // StepLength::MehrotraPredictor (sdpa_parts.cpp:307-354)
        alphaP = alphaD = 0.9;
        if (phase == noINFO || phase == dFEAS) {
            if (alphaP > 1.0) alphaP = 1.0;
        } else {
            dd_real inc = inner(data.C, nw.DxMat);
            if (inc > 0.0) {
                if (alphaP > alphaD) alphaP = alphaD;
                if (alphaP > 1.0) alphaP = 1.0;
            }
        }
        if (phase == noINFO || phase == pFEAS) {
            if (alphaD > 1.0) alphaD = 1.0;
        } else {
            dd_real inc = inner(data.b, nw.DyVec);
            if (inc < 0.0) {
                if (alphaD > alphaP) alphaD = alphaP;
                if (alphaD > 1.0) alphaD = 1.0;
            }
        }
        // DirectionParameter::MehrotraCorrector (sdpa_parts.cpp:499-529)
        {
            dd_real xv = inner(pt.xMat, nw.DzMat), uz = inner(nw.DxMat, pt.zMat), uv = inner(nw.DxMat, nw.DzMat);
            dd_real muTarget = muCurrent + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nDim;
            betaV = muTarget / muCurrent;
            if (betaV < 1.0) betaV = betaV * betaV;
            if (phase == pdFEAS) {
                if (betaV < param.betaStar) betaV = param.betaStar;
                if (betaV > 1.0) betaV = 1.0;
            } else {
                if (betaV < param.betaBar) betaV = param.betaBar;
            }
        }
        if ((phase == pdFEAS) && ((betaV > 5) || (objValPrimal < objValDual))) break;
        mehrotra(false);
        // StepLength::MehrotraCorrector (sdpa_parts.cpp:356-466)
        {
            const dd_real xi = 3.0;
            dd_real alphaBD = 100.0;
            bool eig_ok = true;
            dd_real mx = min_eigen(pt.invCholeskyX, nw.DxMat, eig_ok);
            alphaP = (-mx > 1.0 / alphaBD) ? -1.0 / mx : alphaBD;
            dd_real mz = min_eigen(pt.invCholeskyZ, nw.DzMat, eig_ok);
            if (!eig_ok) {
                res.message = "eigenvalue iteration did not converge";
                break;
            }
            alphaD = (-mz > 1.0 / alphaBD) ? -1.0 / mz : alphaBD;
            alphaP = param.gammaStar * alphaP;
            alphaD = param.gammaStar * alphaD;
            if (phase == noINFO || phase == dFEAS) {
                if (alphaP > 1.0) alphaP = 1.0;
            } else {
                dd_real inc = inner(data.C, nw.DxMat);
                if (inc > 0.0) {
                    if (alphaP > alphaD) alphaP = alphaD;
                    if (alphaP > 1.0) alphaP = 1.0;
                }
            }
            if (phase == noINFO || phase == pFEAS) {
                if (alphaD > 1.0) alphaD = 1.0;
            } else {
                dd_real inc = inner(data.b, nw.DyVec);
                if (inc < 0.0) {
                    if (alphaD > alphaP) alphaD = alphaP;
                    if (alphaD > 1.0) alphaD = 1.0;
                }
            }
            if (switchOn && (phase == noINFO || phase == pFEAS || phase == dFEAS)) {
                dd_real xv = inner(pt.xMat, nw.DzMat), uz = inner(nw.DxMat, pt.zMat), uv = inner(nw.DxMat, nw.DzMat);
                dd_real thetaMax = max((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
                dd_real muNew = muCurrent + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nDim;
                while (thetaMax * thetaMax * muInitial > xi * muNew) {
                    dd_real alphaMax = 0.95 * max(alphaP, alphaD);
                    alphaP = min(alphaP, alphaMax);
                    alphaD = min(alphaD, alphaMax);
                    thetaMax = max((1.0 - alphaP) * thetaP, (1.0 - alphaD) * thetaD);
                    muNew = muCurrent + (alphaP * uz + alphaD * xv + alphaP * alphaD * uv) / nDim;
                    if (alphaP < 1.0e-6 && alphaD < 1.0e-6) break;
                }
            }
            if (phase == pdFEAS) {
                dd_real oD = inner(data.b, pt.yVec);
                dd_real oP = inner(data.C, pt.xMat);
                dd_real incD = inner(data.b, nw.DyVec);
                incD *= alphaD;
                dd_real incP = inner(data.C, nw.DxMat);
                incP *= alphaP;
                dd_real maxRatio = (oD - oP) / (incP - incD);
                if ((maxRatio > 0.0) && (maxRatio < 1.0)) {
                    alphaP *= maxRatio;
                    alphaD *= maxRatio;
                }
            }
        }
        res.trace.push_back({pIteration, muCurrent, thetaP, thetaD, objValPrimal, objValDual, alphaP, alphaD, betaV,
                             std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count()});
        // Solutions::update (sdpa_dataset.cpp:120-143)
        plus(pt.xMat, pt.xMat, nw.DxMat, alphaP);
        Raxpy(m, alphaD, nw.DyVec.data(), 1, pt.yVec.data(), 1);
        plus(pt.zMat, pt.zMat, nw.DzMat, alphaD);
        const dd_real cannot_move = 1.0e-4;
        if (alphaP < cannot_move && alphaD < cannot_move) {
            res.message = "cannot move";
            break;
        }
        if (!compute_inverse(pt, work)) {
            res.message = "cannot move";
            break;
        }
        // RatioInitResCurrentRes::update (sdpa_parts.cpp:632-637)
        if (switchOn) {
            thetaP = abs((1.0 - alphaP) * thetaP);
            thetaD = abs((1.0 - alphaD) * thetaD);
        }
        // AverageComplementarity::update (sdpa_parts.cpp:587-591)
        muCurrent = inner(pt.xMat, pt.zMat);
        muCurrent /= nDim;
        residuals_compute(data, pt, cur);
        // update_exact (sdpa_parts.cpp:639-651)
        thetaP = (initRes.normPrimalVec > 1.0e-10) ? cur.normPrimalVec / initRes.normPrimalVec : dd_real(0.0);
        thetaD = (initRes.normDualMat > 1.0e-10) ? cur.normDualMat / initRes.normDualMat : dd_real(0.0);
        // SolveInfo::update (sdpa_parts.cpp:743-790; with an initial point 691-740)
        {
            objValPrimal = inner(data.C, pt.xMat);
            objValDual = inner(data.b, pt.yVec);
            dd_real primal = thetaP, dual = thetaD, omega = param.omegaStar;
            rho = 0.0;
            dd_real x0z0 = nDim * muInitial, xMatzMat = nDim * muCurrent, x0zMat = 0.0, xMatz0 = 0.0;
            double lam = param.lambdaStar;
            if (init) {  // the initial-point form (sdpa_parts.cpp:691-706)
                x0zMat = inner(initX, pt.zMat);
                xMatz0 = inner(pt.xMat, initZ);
            } else {
                for (size_t b = 0; b < bs.size(); b++) {
                    int dim = bs[b];
                    for (int i = 0; i < dim; i++) {
                        x0zMat += lam * pt.zMat.SDP_block[b].de_ele[(size_t)i * dim + i];
                        xMatz0 += lam * pt.xMat.SDP_block[b].de_ele[(size_t)i * dim + i];
                    }
                }
                for (int b = 0; b < data.LP_nBlock; b++) {
                    x0zMat += lam * pt.zMat.LP_block[b];
                    xMatz0 += lam * pt.xMat.LP_block[b];
                }
            }
            dd_real accuracy = param.epsilonDash;
            if (cur.normPrimalVec <= accuracy && xMatz0 < etaPrimal) etaPrimal = xMatz0;
            if (cur.normDualMat <= accuracy && x0zMat < etaDual) etaDual = x0zMat;
            if (cur.normPrimalVec > accuracy && cur.normDualMat <= accuracy)
                rho = primal * x0zMat / ((primal + (1.0 - primal) * omega) * etaDual + xMatzMat);
            if (cur.normPrimalVec <= accuracy && cur.normDualMat > accuracy)
                rho = dual * xMatz0 / ((dual + (1.0 - dual) * omega) * etaPrimal + xMatzMat);
            if (cur.normPrimalVec > accuracy && cur.normDualMat > accuracy)
                rho = (dual * xMatz0 + primal * x0zMat) /
                      ((primal * dual + omega * (primal * (1.0 - dual) + (1.0 - primal) * dual)) * x0z0 + xMatzMat);
        }
        pIteration++;
    }
    residuals_compute(data, pt, cur);
    res.iterations = pIteration;
    res.phase = phase;
    // Phase::reverse (sdpa_parts.cpp:956-1000)
    int rep = phase;
    switch (phase) {
    case pFEAS: rep = dFEAS; break;
    case dFEAS: rep = pFEAS; break;
    case pFEAS_dINF: rep = pINF_dFEAS; break;
    case pINF_dFEAS: rep = pFEAS_dINF; break;
    case pUNBD: rep = dUNBD; break;
    case dUNBD: rep = pUNBD; break;
    default: break;
    }
    res.phase_reported = rep;
    res.mu = muCurrent;
    res.objValPrimal = objValPrimal;
    res.objValDual = objValDual;
    res.normPrimalVec = cur.normPrimalVec;
    res.normDualMat = cur.normDualMat;
    {
        dd_real mean = (abs(objValPrimal) + abs(objValDual)) / 2.0;
        dd_real PDgap = abs(objValPrimal - objValDual);
        res.relgap = (mean < 1.0) ? PDgap : PDgap / mean;
    }
    res.xMat.clear();
    res.zMat.clear();
    for (size_t l = 0; l < bs.size(); l++) {
        res.xMat.insert(res.xMat.end(), pt.xMat.SDP_block[l].de_ele.begin(), pt.xMat.SDP_block[l].de_ele.end());
        res.zMat.insert(res.zMat.end(), pt.zMat.SDP_block[l].de_ele.begin(), pt.zMat.SDP_block[l].de_ele.end());
    }
    res.xMat.insert(res.xMat.end(), pt.xMat.LP_block.begin(), pt.xMat.LP_block.end());
    res.zMat.insert(res.zMat.end(), pt.zMat.LP_block.begin(), pt.zMat.LP_block.end());
    res.yVec = pt.yVec;
    return true;
}

}  // namespace orc
