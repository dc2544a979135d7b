// mvm_hip.cpp -- H-MVM benchmark driver over hlr::hip::mul_vec.
//
// Follows the reference's mvm program (programs/mvm/mvm.hh:140-212): build
// the matrix, compute y_ref = (50 * 2) A x once, then time samples of 50
// calls y += 2 A x (y reset between samples), report min / median / max
// runtime, GFlop/s and GB/s from the flop and byte model, and the relative
// error of y against y_ref.  x = 1 as in the reference harness
// (programs/mvm/mvm.hh:166,178) unless --xrand is given.
//
// usage: mvm-hip [--app log|logkernel|materncov|invr] [-n N] [--ntile K]
//                [--eta E] [--eps E] [--nbench B] [--prec fp64|fp32|valr|bfp]
//                [--tlr] [--xrand] [--op normal|transposed]
//                [--save FILE | --load FILE]
// --save stores the generated matrix as an H-matrix file; --load reads one
// instead of generating (the problem options then only label the output,
// except --eps, which sets the plan's tolerance).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hlr_hip.hh"
#include "hlrgen.h"

using namespace hlr::hip;

int main(int argc, char** argv) {
    hlrgen_config cfg;
    hlrgen_default_config(&cfg);
    std::string app = "log", prec = "fp64", opname = "normal", save, load;
    int nbench = 5;
    bool xrand = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", a.c_str()); std::exit(2); }
            return argv[++i];
        };
        if (a == "--app") app = next();
        else if (a == "-n") cfg.n = std::atoll(next());
        else if (a == "--ntile") cfg.ntile = std::atoi(next());
        else if (a == "--eta") cfg.eta = std::atof(next());
        else if (a == "--eps") cfg.eps = std::atof(next());
        else if (a == "--nbench") nbench = std::atoi(next());
        else if (a == "--prec") prec = next();
        else if (a == "--op") opname = next();
        else if (a == "--tlr") cfg.format = HLRGEN_FORMAT_TLR;
        else if (a == "--xrand") xrand = true;
        else if (a == "--save") save = next();
        else if (a == "--load") load = next();
        else { std::fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
    }
    if (app == "log") { cfg.kernel = HLRGEN_KERNEL_LOG; cfg.geometry = HLRGEN_GEOM_LINE; }
    else if (app == "logkernel") { cfg.kernel = HLRGEN_KERNEL_LOG_GALERKIN; cfg.geometry = HLRGEN_GEOM_GRID1D; }
    else if (app == "materncov") { cfg.kernel = HLRGEN_KERNEL_MATERN; cfg.geometry = HLRGEN_GEOM_CUBE; }
    else if (app == "invr") { cfg.kernel = HLRGEN_KERNEL_INVR; cfg.geometry = HLRGEN_GEOM_SPHERE; }
    else { std::fprintf(stderr, "unknown app %s\n", app.c_str()); return 2; }
    const matop_t op = (opname == "transposed") ? apply_transposed : apply_normal;

    try {
        int64_t n = 0, ncols = 0;
        std::vector<hmvm_block> blocks;
        std::vector<double> arena;
        if (!load.empty()) {
            const auto t0 = std::chrono::steady_clock::now();
            io::read(load, blocks, arena, n, ncols);
            if (ncols != n) throw std::runtime_error("mvm-hip: the stored matrix is not square");
            std::printf("problem   : %s, n = %lld\n", load.c_str(), (long long)n);
            std::printf("read      : %lld leaves in %.2f s\n", (long long)blocks.size(),
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        } else {
            hlrgen_matrix* gm = nullptr;
            check(hlrgen_build(&cfg, &gm), "hlrgen_build");
            n = ncols = hlrgen_n(gm);
            const int64_t nb = hlrgen_nblocks(gm);
            blocks.assign(hlrgen_blocks(gm), hlrgen_blocks(gm) + nb);
            arena.assign(hlrgen_arena(gm), hlrgen_arena(gm) + hlrgen_arena_len(gm));
            std::printf("problem   : %s, n = %lld, ntile = %d, eta = %g, eps = %g, %s\n", app.c_str(),
                        (long long)n, cfg.ntile, cfg.eta, cfg.eps, cfg.format == HLRGEN_FORMAT_TLR ? "TLR" : "H");
            std::printf("build     : %lld leaves in %.2f s\n", (long long)nb, hlrgen_build_seconds(gm));
            hlrgen_free(gm);
        }
        if (!save.empty()) {
            io::write(blocks, arena, n, n, save);
            std::printf("saved     : %s\n", save.c_str());
        }

        hmvm_plan_opts opts{};
        opts.precision = prec == "fp32" ? HMVM_PREC_FP32 : prec == "valr" ? HMVM_PREC_VALR
                       : prec == "bfp" ? HMVM_PREC_BFP : HMVM_PREC_FP64;
        opts.eps = cfg.eps;
        opts.dense_prec = HMVM_PREC_VALR;          // dense leaves by eps (valr: FP32 unless eps needs FP64)
        opts.check_finite = 1;
        hmatrix A(blocks, arena, n, n, &opts, 0);
        std::printf("plan      : %s, %.3f GB of factors, %lld dense / %lld low-rank leaves\n", prec.c_str(),
                    A.info().factor_bytes * 1e-9, (long long)A.info().ndense, (long long)A.info().nlowrank);

        const int nmvm = 50;
        std::vector<double> xh(n, 1.0);
        if (xrand) {
            hlrgen_random_vector(2, n, -1.0, 1.0, 0, xh.data());
        }
        device_vector<double> x(n), y(n), yref(n);
        x.copy_from(xh);
        hip_arithmetic arith;
        arith.prod(double(nmvm) * 2.0, op, A, x, yref);
        check_hip(hipDeviceSynchronize(), "sync");

        std::vector<double> runtime;
        for (int s = 0; s < nbench; ++s) {
            y.fill_zero();
            check_hip(hipDeviceSynchronize(), "sync");
            const auto t0 = std::chrono::steady_clock::now();
            for (int j = 0; j < nmvm; ++j) mul_vec(2.0, op, A, x, y);
            check_hip(hipDeviceSynchronize(), "sync");
            runtime.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        }
        std::vector<double> sorted = runtime;
        std::sort(sorted.begin(), sorted.end());
        const double tmin = sorted.front(), tmed = sorted[sorted.size() / 2], tmax = sorted.back();
        const double flops = nmvm * A.flops(), bytes = nmvm * A.data_byte_size();
        std::printf("mat-vec   : %d x %d MVMs\n", nbench, nmvm);
        std::printf("  runtime = %.4f s / %.4f s / %.4f s (min / median / max)\n", tmin, tmed, tmax);
        std::printf("  flops   = %.2f GFlop/s\n", flops / tmin * 1e-9);
        std::printf("  bandw.  = %.2f GB/s\n", bytes / tmin * 1e-9);
        std::printf("  MVM/s   = %.1f\n", nmvm / tmin);
        const std::vector<double> a = y.to_host(), b = yref.to_host();
        double num = 0, den = 0;
        for (int64_t i = 0; i < n; ++i) { num += (a[i] - b[i]) * (a[i] - b[i]); den += b[i] * b[i]; }
        std::printf("  error   = %.4e (relative)\n", std::sqrt(num / den));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
