// ORACLE — test infrastructure only.
// Command-line driver: runs the CPU restatement on a .dat-s file and prints
// the result as JSON, every dd value as a [hi, lo] pair with 17 digits.
//   sdpa_oracle <data.dat-s> [param.sdpa] [--dot-threads T] [--max-iter N] [--dump-solution]
//               [--init file [--init-dense]]  (initial point: sparse format unless --init-dense)
//               [--schur auto|dense|sparse]   (the Schur path: the reference's rule, or forced)
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "sdpa_oracle.hpp"

using namespace orc;

static void pdd(const dd_real &v) { std::printf("[%.17e, %.17e]", v.x[0], v.x[1]); }

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s data.dat-s [param] [--dot-threads T] [--max-iter N] [--dump-solution]\n", argv[0]);
        return 2;
    }
    std::string data_path = argv[1], param_path;
    int max_iter = -1;
    bool dump = false, init_dense = false;
    std::string init_path;
    for (int i = 2; i < argc; i++) {
        if (!std::strcmp(argv[i], "--dot-threads") && i + 1 < argc) g_dot_threads = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--max-iter") && i + 1 < argc) max_iter = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--dump-solution")) dump = true;
        else if (!std::strcmp(argv[i], "--init") && i + 1 < argc) init_path = argv[++i];
        else if (!std::strcmp(argv[i], "--init-dense")) init_dense = true;
        else if (!std::strcmp(argv[i], "--schur") && i + 1 < argc) {
            const char *v = argv[++i];
            g_schur_mode = !std::strcmp(v, "dense") ? 1 : !std::strcmp(v, "sparse") ? 2 : 0;
        }
        else param_path = argv[i];
    }
    InputData data;
    std::string err;
    if (!read_dat_s(data_path, data, err)) {
        std::printf("{\"error\": \"%s\"}\n", err.c_str());
        return 1;
    }
    Parameter param;
    if (!param_path.empty() && !param.readFile(param_path)) {
        std::printf("{\"error\": \"cannot read parameter file\"}\n");
        return 1;
    }
    if (max_iter >= 0) param.maxIteration = max_iter;
    InitPoint ip;
    if (!init_path.empty() && !read_init(init_path, data, param, !init_dense, ip, err)) {
        std::printf("{\"error\": \"%s\"}\n", err.c_str());
        return 1;
    }
    SolveResult res;
    auto t0 = std::chrono::steady_clock::now();
    solve(data, param, res, init_path.empty() ? nullptr : &ip);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    static const char *names[] = {"noINFO", "pFEAS", "dFEAS", "pdFEAS", "pdINF", "pFEAS_dINF", "pINF_dFEAS", "pdOPT", "pUNBD", "dUNBD"};
    std::printf("{\"m\": %d, \"iterations\": %d, \"phase\": \"%s\", \"phase_internal\": \"%s\", \"seconds\": %.6f,\n",
                data.m, res.iterations, names[res.phase_reported], names[res.phase], secs);
    std::printf(" \"message\": \"%s\", \"dot_threads\": %d,\n", res.message.c_str(), g_dot_threads);
    std::printf(" \"schur\": \"%s\", \"schur_nnz\": %lld, \"schur_ordering\": [", res.schur_sparse ? "sparse" : "dense",
                res.schur_nnz);
    for (size_t i = 0; i < res.schur_ordering.size(); i++) std::printf("%s%d", i ? ", " : "", res.schur_ordering[i]);
    std::printf("],\n");
    std::printf(" \"mu\": "); pdd(res.mu);
    std::printf(", \"objValPrimal\": "); pdd(res.objValPrimal);
    std::printf(", \"objValDual\": "); pdd(res.objValDual);
    std::printf(",\n \"normPrimalVec\": "); pdd(res.normPrimalVec);
    std::printf(", \"normDualMat\": "); pdd(res.normDualMat);
    std::printf(", \"relgap\": "); pdd(res.relgap);
    std::printf(",\n \"trace\": [");
    for (size_t i = 0; i < res.trace.size(); i++) {
        const IterRecord &r = res.trace[i];
        std::printf("%s\n  {\"it\": %d, \"mu\": ", i ? "," : "", r.it); pdd(r.mu);
        std::printf(", \"thetaP\": "); pdd(r.thetaP);
        std::printf(", \"thetaD\": "); pdd(r.thetaD);
        std::printf(", \"objP\": "); pdd(r.objP);
        std::printf(", \"objD\": "); pdd(r.objD);
        std::printf(", \"alphaP\": "); pdd(r.alphaP);
        std::printf(", \"alphaD\": "); pdd(r.alphaD);
        std::printf(", \"beta\": "); pdd(r.beta);
        std::printf(", \"seconds\": %.6f}", r.seconds);
    }
    std::printf("]");
    if (dump) {
        auto arr = [](const char *name, const std::vector<dd_real> &v) {
            std::printf(",\n \"%s\": [", name);
            for (size_t i = 0; i < v.size(); i++) {
                if (i) std::printf(", ");
                pdd(v[i]);
            }
            std::printf("]");
        };
        arr("yVec", res.yVec);
        arr("xMat", res.xMat);
        arr("zMat", res.zMat);
    }
    std::printf("}\n");
    return 0;
}
