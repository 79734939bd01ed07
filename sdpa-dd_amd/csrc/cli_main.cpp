// sdpa_dd_gpu: command-line front end in the shape of the reference's
// sdpa_dd (sdpa_main.cpp:523-632): -ds data.dat-s [-p param] [-o out]
// [-is init.ini-s | -id init.ini].  Prints the result as JSON (dd values as
// [hi, lo] pairs) on stdout and, with -o, the reference-style final block to
// the output file; --dump file writes the final point as a full-dd sparse
// initial-point file.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sdpadd.h"

static void pdd(FILE *f, const double *v) { std::fprintf(f, "[%.17e, %.17e]", v[0], v[1]); }

int main(int argc, char **argv) {
    std::string data, param, outp, initf, dump;
    bool init_sparse = false;
    int dot_chunks = 1024;
    bool trace = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "-ds") && i + 1 < argc) data = argv[++i];
        else if (!std::strcmp(argv[i], "-p") && i + 1 < argc) param = argv[++i];
        else if (!std::strcmp(argv[i], "-o") && i + 1 < argc) outp = argv[++i];
        else if (!std::strcmp(argv[i], "-is") && i + 1 < argc) initf = argv[++i], init_sparse = true;
        else if (!std::strcmp(argv[i], "-id") && i + 1 < argc) initf = argv[++i], init_sparse = false;
        else if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
        else if (!std::strcmp(argv[i], "--dot-partition") && i + 1 < argc) dot_chunks = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--trace")) trace = true;
    }
    if (data.empty()) {
        std::fprintf(stderr,
                     "usage: %s -ds data.dat-s [-p param.sdpa] [-o out] [-is init.ini-s | -id init.ini] [--dump point]"
                     " [--dot-partition T] [--trace]\n",
                     argv[0]);
        return 2;
    }
    sdpadd_handle h;
    if (sdpadd_create(0, &h) != 0) {
        std::fprintf(stderr, "no GPU\n");
        return 1;
    }
    sdpadd_set_dot_partition(h, dot_chunks);
    sdpadd_params prm;
    sdpadd_default_params(&prm);
    if (!param.empty() && sdpadd_read_params(param.c_str(), &prm) != 0) {
        std::fprintf(stderr, "cannot read %s\n", param.c_str());
        return 1;
    }
    sdpadd_problem p;
    if (sdpadd_problem_load(h, data.c_str(), &p) != 0) {
        std::printf("{\"error\": \"%s\"}\n", sdpadd_last_error(h));
        return 1;
    }
    if (!initf.empty() && sdpadd_read_initial_point(p, initf.c_str(), init_sparse) != 0) {
        std::printf("{\"error\": \"%s\"}\n", sdpadd_last_error(h));
        return 1;
    }
    sdpadd_result r;
    int rc = sdpadd_solve(p, &prm, &r);
    if (rc != 0) {
        std::printf("{\"error\": \"solve failed: %s\"}\n", sdpadd_last_error(h));
        return 1;
    }
    static const char *names[] = {"noINFO", "pFEAS", "dFEAS", "pdFEAS", "pdINF", "pFEAS_dINF", "pINF_dFEAS", "pdOPT", "pUNBD", "dUNBD"};
    std::printf("{\"iterations\": %d, \"phase\": \"%s\", \"phase_internal\": \"%s\", \"seconds\": %.6f,\n", r.iterations,
                names[r.phase], names[r.phase_internal], r.seconds);
    std::printf(" \"t_schur\": %.6f, \"t_chol\": %.6f, \"t_lanczos\": %.6f, \"t_inverse\": %.6f,\n", r.t_schur, r.t_chol,
                r.t_lanczos, r.t_inverse);
    std::printf(" \"mu\": "); pdd(stdout, r.mu);
    std::printf(", \"objValPrimal\": "); pdd(stdout, r.objValPrimal);
    std::printf(", \"objValDual\": "); pdd(stdout, r.objValDual);
    std::printf(",\n \"normPrimalVec\": "); pdd(stdout, r.normPrimalVec);
    std::printf(", \"normDualMat\": "); pdd(stdout, r.normDualMat);
    std::printf(", \"relgap\": "); pdd(stdout, r.relgap);
    if (trace) {
        std::vector<double> t(17 * 1000);
        int n = sdpadd_trace(p, t.data(), 1000);
        std::printf(",\n \"trace\": [");
        const char *keys[] = {"mu", "thetaP", "thetaD", "objP", "objD", "alphaP", "alphaD", "beta"};
        for (int i = 0; i < n && i < 1000; i++) {
            std::printf("%s\n  {\"it\": %d", i ? "," : "", (int)t[17 * i]);
            for (int q = 0; q < 8; q++) {
                std::printf(", \"%s\": ", keys[q]);
                pdd(stdout, &t[17 * i + 1 + 2 * q]);
            }
            std::printf("}");
        }
        std::printf("]");
    }
    std::printf("}\n");
    if (!outp.empty()) {
        FILE *f = std::fopen(outp.c_str(), "w");
        if (f) {
            // REVERSE_PRIMAL_DUAL printing (sdpa_io.cpp:1100-1121)
            std::fprintf(f, "phase.value = %s\n", names[r.phase]);
            std::fprintf(f, "   Iteration = %d\n", r.iterations);
            std::fprintf(f, "          mu = %4.16e\n", r.mu[0]);
            std::fprintf(f, "relative gap = %4.16e\n", r.relgap[0]);
            std::fprintf(f, "objValPrimal = %10.16e\n", -r.objValDual[0]);
            std::fprintf(f, "objValDual   = %10.16e\n", -r.objValPrimal[0]);
            std::fprintf(f, "p.feas.error = %10.16e\n", r.normDualMat[0]);
            std::fprintf(f, "d.feas.error = %10.16e\n", r.normPrimalVec[0]);
            std::fprintf(f, "    main loop time = %.6f\n", r.seconds);
            std::fclose(f);
        }
    }
    if (!dump.empty() && sdpadd_write_solution(p, dump.c_str()) != 0) {
        std::fprintf(stderr, "%s\n", sdpadd_last_error(h));
        return 1;
    }
    sdpadd_problem_destroy(p);
    sdpadd_destroy(h);
    return 0;
}
