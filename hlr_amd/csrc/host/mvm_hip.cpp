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
//                [--save FILE | --load FILE] [--norm]
// --save stores the generated matrix as an H-matrix file; --load reads one
// instead of generating (the problem options then only label the output,
// except --eps, which sets the plan's tolerance).
// --norm: the spectral norm twice, once by a generic power iteration written
// against HLR's arithmetic concept and run with hip_arithmetic on host
// vectors (the plugin path), once by hlr::hip::norm::spectral on the device.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "hlr_hip.hh"
#include "hlrgen.h"

// HLR's arithmetic trait and concept (arith/defaults.hh:20-24), defined after
// hlr_hip.hh has declared the trait and registered hip_arithmetic with it
namespace hlr {
template <typename T> struct is_arithmetic { static constexpr bool value = false; };
template <typename T> inline constexpr bool is_arithmetic_v = is_arithmetic<T>::value;
template <typename T> concept arithmetic_type = is_arithmetic_v<T>;
}

using namespace hlr::hip;

static_assert(hlr::arithmetic_type<hip_arithmetic> && hlr::arithmetic_type<const hip_arithmetic&>,
              "hip_arithmetic satisfies HLR's arithmetic concept");

namespace {

// A generic algorithm in the shape of HLR's norm::spectral
// (arith/detail/norm.hh:448-527): constrained by arithmetic_type, it sees
// the operator only through nrows/ncols and arithmetic.prod on host vectors.
template <hlr::arithmetic_type arithmetic_t, typename operator_t>
double generic_spectral(arithmetic_t&& arithmetic, const operator_t& A, const double atol, const size_t amax_it) {
    const size_t nr = nrows(A), nc = ncols(A);
    std::vector<double> xv(nc), yv(nc), tv(nr);
    host_vector_ref<double> x{xv}, y{yv}, t{tv};
    auto gen = std::default_random_engine();
    auto dist = std::uniform_real_distribution<double>(-1.0, 1.0);
    for (auto& v : xv) v = dist(gen);
    auto norm2 = [](const std::vector<double>& v) { double s = 0; for (double e : v) s += e * e; return std::sqrt(s); };
    const double nx = norm2(xv);
    for (auto& v : xv) v /= nx;
    const size_t max_it = amax_it == 0 ? std::max(size_t(5), std::min(nr, nc) / 10) : amax_it;
    const double eps = std::numeric_limits<double>::epsilon();
    const double tol = atol == 0 ? std::sqrt(eps) : atol, abs_tol = std::min(10 * eps, tol), zero = eps * eps;
    double lambda = 1.0;
    for (size_t i = 0; i < max_it; ++i) {
        std::fill(tv.begin(), tv.end(), 0.0);
        arithmetic.prod(1.0, apply_normal, A, x, t);
        std::fill(yv.begin(), yv.end(), 0.0);
        arithmetic.prod(1.0, apply_adjoint, A, t, y);
        double dot = 0;
        for (size_t j = 0; j < nc; ++j) dot += xv[j] * yv[j];
        const double lambda_new = std::sqrt(std::fabs(dot)), norm_y = norm2(yv);
        if (std::fabs((lambda_new - lambda) / lambda) < tol) return lambda_new;
        if (i > 5 && std::fabs(lambda_new - lambda) < abs_tol) return lambda_new;
        if (lambda_new < zero) return lambda_new;
        lambda = lambda_new;
        if (norm_y <= zero) break;
        for (size_t j = 0; j < nc; ++j) xv[j] = yv[j] / norm_y;
    }
    return lambda;
}

} // namespace

int main(int argc, char** argv) {
    hlrgen_config cfg;
    hlrgen_default_config(&cfg);
    std::string app = "log", prec = "fp64", opname = "normal", save, load;
    int nbench = 5;
    bool xrand = false, norm_check = false;
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
        else if (a == "--norm") norm_check = true;
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

        if (norm_check) {
            hip_arithmetic arith;
            const double gen = generic_spectral(arith, A, 1e-10, 40);
            const double dev = hlr::hip::norm::spectral(arith, A, 1e-10, 40);
            std::printf("norm      : generic %.15e device %.15e rel.diff %.3e\n", gen, dev,
                        std::fabs(gen - dev) / dev);
            return 0;
        }

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
