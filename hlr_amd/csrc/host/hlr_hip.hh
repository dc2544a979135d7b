// hlr_hip.hh -- C++ mirror of HLR's mul_vec call surface over the hmvm C-ABI.
//
// Header-only.  Gives C++ callers the reference's spelling:
//   hlr::hip::mul_vec(alpha, op, M, x, y)      (tbb/arith.hh:312-340,
//                                               seq/arith.hh:40-72)
//   hlr::hip::tlr::mul_vec(alpha, op, M, x, y) (tbb/arith.hh:998-1042)
//   hlr::hip::hip_arithmetic                   (tbb/arith.hh:1414-1510,
//                                               arith/defaults.hh:20-104)
//   hlr::hip::norm::spectral(arith, A, ...)    (arith/norm.hh:89-101)
//   hlr::hip::norm::inv_error_2(arith, A, W)   (arith/norm.hh:115-137)
//   hlr::hip::linear_iteration(A, W, b, x, ...) (programs/iterrefine.hh:20-80)
//   hlr::hip::io::write / read                 (utils/io.hh:59-80)
// with y += alpha op(M) x, alpha == 0 a no-op, op in {apply_normal,
// apply_conjugate, apply_transposed, apply_adjoint}, and errors thrown as
// std::runtime_error (HLR_ASSERT/HLR_ERROR, utils/log.hh:40-67).  The matrix
// is an hmvm plan on one HIP device; vectors are device arrays in the
// internal (cluster) ordering.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hmvm.h"

namespace hlr {

// HLR's arithmetic trait (arith/defaults.hh:20-24: primary template false,
// `arithmetic_type` is the concept is_arithmetic_v<T>).  Declared here, not
// defined, so that hip_arithmetic registers with it below whether HLR's
// headers are included before this one, after it, or not at all.
template <typename T> struct is_arithmetic;

namespace hip {

enum matop_t { apply_normal = HMVM_OP_NORMAL, apply_transposed = HMVM_OP_TRANSPOSED,
               apply_adjoint = HMVM_OP_ADJOINT, apply_conjugate = HMVM_OP_CONJUGATE };

inline void check(int rc, const char* what) {
    if (rc != HMVM_OK) throw std::runtime_error(std::string(what) + ": " + hmvm_error_string(rc));
}

inline void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// device-resident vector of value_t (double or float) of length n, possibly
// with several columns (nrhs) of leading dimension n
template <typename value_t>
class device_vector {
public:
    device_vector() = default;
    explicit device_vector(size_t n, int ncols = 1) : n_(n), ncols_(ncols) {
        check_hip(hipMalloc(&p_, sizeof(value_t) * n * size_t(ncols)), "hipMalloc");
        check_hip(hipMemset(p_, 0, sizeof(value_t) * n * size_t(ncols)), "hipMemset");
    }
    device_vector(const device_vector&) = delete;
    device_vector& operator=(const device_vector&) = delete;
    device_vector(device_vector&& o) noexcept { swap(o); }
    device_vector& operator=(device_vector&& o) noexcept { swap(o); return *this; }
    ~device_vector() { if (p_) (void)hipFree(p_); }

    void copy_from(const std::vector<value_t>& h) {
        if (h.size() != n_ * size_t(ncols_)) throw std::runtime_error("device_vector: size mismatch");
        check_hip(hipMemcpy(p_, h.data(), sizeof(value_t) * h.size(), hipMemcpyHostToDevice), "hipMemcpy");
    }
    std::vector<value_t> to_host() const {
        std::vector<value_t> h(n_ * size_t(ncols_));
        check_hip(hipMemcpy(h.data(), p_, sizeof(value_t) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
        return h;
    }
    void fill_zero() { check_hip(hipMemset(p_, 0, sizeof(value_t) * n_ * size_t(ncols_)), "hipMemset"); }

    value_t* data() { return p_; }
    const value_t* data() const { return p_; }
    size_t length() const { return n_; }
    int ncols() const { return ncols_; }

private:
    void swap(device_vector& o) { std::swap(p_, o.p_); std::swap(n_, o.n_); std::swap(ncols_, o.ncols_); }
    value_t* p_ = nullptr;
    size_t n_ = 0;
    int ncols_ = 1;
};

// H- or TLR-matrix held as a device plan
class hmatrix {
public:
    hmatrix(const std::vector<hmvm_block>& blocks, const std::vector<double>& arena, int64_t nrows,
            int64_t ncols, const hmvm_plan_opts* opts = nullptr, int device = 0) {
        check(hmvm_plan_create(blocks.data(), int64_t(blocks.size()), nrows, ncols, arena.data(),
                               int64_t(arena.size()), opts, device, &p_),
              "hmvm_plan_create");
        check(hmvm_plan_get_info(p_, &info_), "hmvm_plan_get_info");
    }
    // from an H-matrix file written by io::write
    explicit hmatrix(const std::string& filename, const hmvm_plan_opts* opts = nullptr, int device = 0) {
        check(hmvm_plan_create_from_file(filename.c_str(), opts, device, &p_), "hmvm_plan_create_from_file");
        check(hmvm_plan_get_info(p_, &info_), "hmvm_plan_get_info");
    }
    hmatrix(const hmatrix&) = delete;
    hmatrix& operator=(const hmatrix&) = delete;
    ~hmatrix() { if (p_) hmvm_plan_destroy(p_); }

    using value_t = double;       // factors enter as FP64 (the plan may store them narrower)

    const hmvm_plan* plan() const { return p_; }
    int64_t nrows() const { return info_.nrows; }
    int64_t ncols() const { return info_.ncols; }
    int64_t nrows(matop_t op) const { return (op == apply_transposed || op == apply_adjoint) ? ncols() : nrows(); }
    int64_t ncols(matop_t op) const { return (op == apply_transposed || op == apply_adjoint) ? nrows() : ncols(); }
    const hmvm_plan_info& info() const { return info_; }
    // mul_vec_flops / mul_vec_datasize (arith/mulvec.hh:795-874)
    double flops() const { return info_.flops; }
    double data_byte_size() const { return info_.datasize; }

private:
    hmvm_plan* p_ = nullptr;
    hmvm_plan_info info_{};
};

template <typename value_t> constexpr int vec_dtype();
template <> constexpr int vec_dtype<double>() { return HMVM_F64; }
template <> constexpr int vec_dtype<float>() { return HMVM_F32; }

// y += alpha op(M) x
template <typename value_t>
void mul_vec(const value_t alpha, const matop_t op, const hmatrix& M, const device_vector<value_t>& x,
             device_vector<value_t>& y, hipStream_t stream = nullptr) {
    if (int64_t(x.length()) != M.ncols(op) || int64_t(y.length()) != M.nrows(op) || x.ncols() != y.ncols())
        throw std::runtime_error("mul_vec: vector sizes do not match the matrix");
    check(hmvm_apply(M.plan(), double(alpha), int(op), x.data(), int64_t(x.length()), y.data(),
                     int64_t(y.length()), x.ncols(), vec_dtype<value_t>(), stream),
          "hmvm_apply");
}

// H-matrix files: the flattened leaves and FP64 arena that hmatrix takes
// (in place of io::hpro::write / read, utils/io.hh:59-80)
namespace io {

inline void write(const std::vector<hmvm_block>& blocks, const std::vector<double>& arena, int64_t nrows,
                  int64_t ncols, const std::string& filename) {
    check(hmvm_hmatrix_write(filename.c_str(), blocks.data(), int64_t(blocks.size()), nrows, ncols, arena.data(),
                             int64_t(arena.size())),
          "hmvm_hmatrix_write");
}

inline void read(const std::string& filename, std::vector<hmvm_block>& blocks, std::vector<double>& arena,
                 int64_t& nrows, int64_t& ncols) {
    int64_t nb = 0, na = 0;
    check(hmvm_hmatrix_read_info(filename.c_str(), &nrows, &ncols, &nb, &na), "hmvm_hmatrix_read_info");
    blocks.resize(size_t(nb));
    arena.resize(size_t(na));
    check(hmvm_hmatrix_read(filename.c_str(), blocks.data(), nb, arena.data(), na), "hmvm_hmatrix_read");
}

} // namespace io

namespace tlr {

// y += alpha op(M) x for a one-level (tile-low-rank) matrix
template <typename value_t>
void mul_vec(const value_t alpha, const matop_t op, const hmatrix& M, const device_vector<value_t>& x,
             device_vector<value_t>& y, hipStream_t stream = nullptr) {
    if (int64_t(x.length()) != M.ncols(op) || int64_t(y.length()) != M.nrows(op))
        throw std::runtime_error("tlr::mul_vec: vector sizes do not match the matrix");
    check(hmvm_tlr_apply(M.plan(), double(alpha), int(op), x.data(), y.data(), vec_dtype<value_t>(), stream),
          "hmvm_tlr_apply");
}

} // namespace tlr

// sizes of an operator, as generic algorithms ask for them (nrows(A),
// ncols(A) in arith/detail/norm.hh:458-459), found by argument lookup
inline size_t nrows(const hmatrix& M) { return size_t(M.nrows()); }
inline size_t ncols(const hmatrix& M) { return size_t(M.ncols()); }

// arithmetic object for generic algorithms (norm::spectral and friends call
// arithmetic.prod(alpha, op, A, x, y), arith/detail/norm.hh:492-494).
// Registered with hlr::is_arithmetic below, so it satisfies HLR's
// `arithmetic_type` concept.  Two kinds of vectors:
//   - device_vector: stream-ordered on `stream`, as hlr::hip::mul_vec;
//   - host vectors with data() and length() (HLR's blas::vector and
//     scalar_vector, std::vector through host_vector_ref): staged through
//     device memory and synchronous, as the reference's call is.
struct hip_arithmetic {
    hipStream_t stream = nullptr;

    template <typename value_t>
    void mul_vec(const value_t alpha, const matop_t op, const hmatrix& M, const device_vector<value_t>& x,
                 device_vector<value_t>& y) const {
        hlr::hip::mul_vec(alpha, op, M, x, y, stream);
    }

    template <typename value_t>
    void prod(const value_t alpha, const matop_t op, const hmatrix& M, const device_vector<value_t>& x,
              device_vector<value_t>& y) const {
        hlr::hip::mul_vec(alpha, op, M, x, y, stream);
    }

    template <typename value_t, typename host_vector_t>
    void mul_vec(const value_t alpha, const matop_t op, const hmatrix& M, const host_vector_t& x,
                 host_vector_t& y) const {
        if (alpha == value_t(0)) return;                     // tbb/arith.hh:54-55: y untouched
        const size_t nx = size_t(x.length()), ny = size_t(y.length());
        if (int64_t(nx) != M.ncols(op) || int64_t(ny) != M.nrows(op))
            throw std::runtime_error("prod: vector sizes do not match the matrix");
        device_vector<value_t> dx(nx), dy(ny);
        check_hip(hipMemcpyAsync(dx.data(), x.data(), sizeof(value_t) * nx, hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync");
        check_hip(hipMemcpyAsync(dy.data(), y.data(), sizeof(value_t) * ny, hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync");
        hlr::hip::mul_vec(alpha, op, M, dx, dy, stream);
        check_hip(hipMemcpyAsync(y.data(), dy.data(), sizeof(value_t) * ny, hipMemcpyDeviceToHost, stream),
                  "hipMemcpyAsync");
        check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    template <typename value_t, typename host_vector_t>
    void prod(const value_t alpha, const matop_t op, const hmatrix& M, const host_vector_t& x,
              host_vector_t& y) const {
        mul_vec(alpha, op, M, x, y);
    }
};

// a std::vector seen through the data() / length() interface of HLR's vectors
template <typename value_t>
struct host_vector_ref {
    std::vector<value_t>& v;
    value_t* data() const { return v.data(); }
    size_t length() const { return v.size(); }
};

namespace norm {

// spectral norm by power iteration on the device (arith/detail/norm.hh:448-527,
// called as hlr::norm::spectral(arithmetic, A, atol, amax_it, squared) with
// the defaults of arith/norm.hh:89-101)
inline double spectral(const hip_arithmetic& arith, const hmatrix& A, const double atol = 1e-3,
                       const size_t amax_it = 50, const bool squared = true) {
    double lambda = 0;
    check(hmvm_spectral_norm(A.plan(), atol, int64_t(amax_it), squared ? 1 : 0, arith.stream, &lambda, nullptr),
          "hmvm_spectral_norm");
    return lambda;
}

// |I - A A_inv|_2 (arith/norm.hh:115-137): norm::spectral with its defaults
// on the operator I - A A_inv, applied factor by factor
inline double inv_error_2(const hip_arithmetic& arith, const hmatrix& A, const hmatrix& A_inv) {
    double err = 0;
    check(hmvm_inv_error_2(A.plan(), A_inv.plan(), 1e-3, 50, arith.stream, &err, nullptr), "hmvm_inv_error_2");
    return err;
}

} // namespace norm

// linear iteration x <- x - W (A x - b) (programs/iterrefine.hh:20-80) until
// |sol - x| <= tol (or |A x - b| <= tol |b| without sol) or max_steps steps;
// returns the number of steps, and the per-step |r|, |c|, |e| in history
inline int64_t linear_iteration(const hmatrix& A, const hmatrix& W, const device_vector<double>& b,
                                device_vector<double>& x, const device_vector<double>* sol = nullptr,
                                const double tol = 1e-8, const int64_t max_steps = 100,
                                std::vector<double>* history = nullptr, hipStream_t stream = nullptr) {
    const size_t n = size_t(A.nrows());
    if (A.ncols() != A.nrows() || W.nrows() != A.nrows() || W.ncols() != A.nrows() || b.length() != n ||
        x.length() != n || (sol && sol->length() != n))
        throw std::runtime_error("linear_iteration: sizes do not match");
    // the library takes max_steps == 0 as 100 steps
    if (history) history->assign(size_t(3 * (max_steps == 0 ? 100 : std::max<int64_t>(max_steps, 1))), 0.0);
    int64_t steps = 0;
    check(hmvm_linear_iteration(A.plan(), W.plan(), b.data(), x.data(), sol ? sol->data() : nullptr, tol,
                                max_steps, stream, &steps, history ? history->data() : nullptr),
          "hmvm_linear_iteration");
    if (history) history->resize(size_t(3 * steps));
    return steps;
}

} // namespace hip

// registration with HLR's trait, for the value and reference forms generic
// algorithms deduce (as arith/defaults.hh:101-104 does for default_arithmetic)
template <> struct is_arithmetic<hip::hip_arithmetic> { static constexpr bool value = true; };
template <> struct is_arithmetic<const hip::hip_arithmetic> { static constexpr bool value = true; };
template <> struct is_arithmetic<hip::hip_arithmetic&> { static constexpr bool value = true; };
template <> struct is_arithmetic<const hip::hip_arithmetic&> { static constexpr bool value = true; };

} // namespace hlr
