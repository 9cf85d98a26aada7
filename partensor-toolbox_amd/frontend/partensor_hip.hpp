// partensor_hip.hpp -- C++ front end of the MI355X execution policy.
//
// Mirrors the reference's policy-dispatched interface:
//   partensor::cpd(partensor::execution::hip, tnsX, R [, options] [, factorsInit])
//   partensor::cpdDimTree(partensor::execution::hip, ...)
// with the file-based forms (dims, R, path [, options]), (dims, R, paths [, options])
// and (dims, R, true_paths, init_paths, options) (Cpd.hpp:893-1600,
// CpdDimTree.hpp:807-1327).  Everything runs through the C ABI of ptn_hip.h; errors are
// raised as std::runtime_error.
//
// Two ways to include it:
//  * Inside a reference build, after the reference's headers (PARTENSOR.hpp).  Then
//    HipOptions<Tensor> is Options<Tensor, execution::hip_policy, DefaultValues> -- the
//    pattern of CudaOptions (PARTENSOR_basic.hpp:771-775) -- HipStatus<Tensor> is that
//    policy's Status<...> plus the per-sweep log, is_execution_policy<hip_policy> is
//    specialised (execution.hpp:283-316), and the reference's Tensor<N> / Matrix are the
//    tensor and matrix types.  The overloads below take execution::hip_policy by value;
//    partial ordering makes them win over the reference's generic cpd(ExecutionPolicy&&, ...)
//    templates, so cpd(execution::hip, tnsX, R, options) reaches the GPU without an edit of
//    Cpd.hpp.  This header defines nothing the reference defines.
//  * On its own (no Eigen, no reference headers).  Then HipOptions / HipStatus are this
//    header's hip::Options / hip::Status with every field of the reference's Options /
//    Status under the same names and defaults (PARTENSOR_basic.hpp:183-209, 283-336,
//    556-595), including method, lambdas and max_duration, which the ALS path does not
//    consume, and hip::Tensor<N> / hip::Matrix are the tensor and matrix types.  If the
//    reference's execution.hpp or Constants.hpp was included first, the policy trait is
//    specialised and the reference's Constraint / Method enums are used.
//
// The MI355X choices (device, contraction arithmetic, fused line-search pass, Nesterov
// cap, seed of the initial factors) travel on the policy object, not in Options, so that
// Options stays the reference's type: partensor::execution::hip is the plain FP64 path on
// device 0; e.g. execution::hip.with_arith(hip::Arith::bf16x3).with_fused_line_search()
// selects the FP32-stream form.
#ifndef PARTENSOR_HIP_HPP
#define PARTENSOR_HIP_HPP

#include "ptn_hip.h"

#include <array>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace partensor {
namespace hip {

// Arithmetic of the tensor contractions (PTN_ARITH_*).
enum class Arith { f64 = PTN_ARITH_F64, f32 = PTN_ARITH_F32, bf16x3 = PTN_ARITH_BF16X3, f32_f64 = PTN_ARITH_F32_F64 };

}  // namespace hip

namespace execution {

// Execution-policy tag in the shape of the reference's policies (execution.hpp:170-203),
// carrying this policy's execution choices.
class hip_policy {
 public:
  static constexpr std::false_type __allow_unsequenced() { return std::false_type{}; }
  static constexpr std::false_type __allow_vector() { return std::false_type{}; }
  static constexpr std::false_type __allow_parallel() { return std::false_type{}; }
  static constexpr std::false_type __allow_cuda() { return std::false_type{}; }
  static constexpr std::false_type __allow_openmpi() { return std::false_type{}; }
  static constexpr std::false_type __allow_openmp() { return std::false_type{}; }
  static constexpr std::true_type __allow_hip() { return std::true_type{}; }

  int device = 0;                       // HIP device ordinal
  hip::Arith arith = hip::Arith::f64;   // tensor contractions; f64 is the plain path
  bool fuse_accel = false;              // line-search pass fused into the next left node (tree, FP32)
  int max_inner = 50;                   // Nesterov cap, NesterovMNLS.hpp:187
  uint64_t init_seed = 1;               // seeded U[0,1) start when no factors are given

  constexpr hip_policy on_device(int d) const { hip_policy p(*this); p.device = d; return p; }
  constexpr hip_policy with_arith(hip::Arith a) const { hip_policy p(*this); p.arith = a; return p; }
  constexpr hip_policy with_fused_line_search(bool on = true) const {
    hip_policy p(*this);
    p.fuse_accel = on;
    return p;
  }
  constexpr hip_policy with_max_inner(int n) const { hip_policy p(*this); p.max_inner = n; return p; }
  constexpr hip_policy with_init_seed(uint64_t s) const { hip_policy p(*this); p.init_seed = s; return p; }
};

constexpr hip_policy hip{};

#ifdef PARTENSOR_EXECUTION_HPP
// the reference's trait (execution.hpp:283-316): hip is an execution policy
template <>
struct is_execution_policy<hip_policy> : std::true_type {};
#endif

}  // namespace execution

namespace hip {

// Column-major dense tensor of doubles, mode 0 fastest (Tensor.hpp:63-64).
template <int N>
class Tensor {
 public:
  static constexpr int NumDimensions = N;
  Tensor() { d_.fill(0); }
  explicit Tensor(const std::array<int64_t, N> &dims) { resize(dims); }
  void resize(const std::array<int64_t, N> &dims) {
    d_ = dims;
    int64_t n = 1;
    for (int64_t x : d_) n *= x;
    v_.assign((size_t)n, 0.0);
  }
  int64_t dimension(int i) const { return d_[i]; }
  const std::array<int64_t, N> &dimensions() const { return d_; }
  int64_t size() const { return (int64_t)v_.size(); }
  double *data() { return v_.data(); }
  const double *data() const { return v_.data(); }
  template <typename... Ix>
  double &operator()(Ix... ix) {
    static_assert(sizeof...(Ix) == N, "one index per mode");
    const int64_t idx[N] = {(int64_t)ix...};
    int64_t e = 0, st = 1;
    for (int k = 0; k < N; ++k) { e += idx[k] * st; st *= d_[k]; }
    return v_[(size_t)e];
  }

 private:
  std::array<int64_t, N> d_;
  std::vector<double> v_;
};

// Column-major dense matrix of doubles (the layout of Eigen::MatrixXd).
class Matrix {
 public:
  Matrix() = default;
  Matrix(int64_t rows, int64_t cols) : r_(rows), c_(cols), v_((size_t)(rows * cols), 0.0) {}
  int64_t rows() const { return r_; }
  int64_t cols() const { return c_; }
  double *data() { return v_.data(); }
  const double *data() const { return v_.data(); }
  double &operator()(int64_t i, int64_t j) { return v_[(size_t)(i + r_ * j)]; }
  double operator()(int64_t i, int64_t j) const { return v_[(size_t)(i + r_ * j)]; }

 private:
  int64_t r_ = 0, c_ = 0;
  std::vector<double> v_;
};

// Tensor type -> order and factor matrix type, for this header's templates.
template <typename T>
struct TensorTraits;

template <int N>
struct TensorTraits<Tensor<N>> {
  static constexpr std::size_t TnsSize = N;
  using MatrixType = Matrix;
};

#ifdef PARTENSOR_TYPES_HPP
// the reference's Eigen tensor and matrix (Tensor.hpp:42-64)
template <int N>
struct TensorTraits<partensor::Tensor<N>> {
  static constexpr std::size_t TnsSize = N;
  using MatrixType = partensor::Matrix;
};
#endif

template <typename Tensor_>
using MatrixArray = std::array<typename TensorTraits<Tensor_>::MatrixType, TensorTraits<Tensor_>::TnsSize>;

#ifdef PARTENSOR_BASIC_HPP
// The tensor type the file forms read into: the reference's when it is present.
template <int N>
using FileTensor = partensor::Tensor<N>;
#else
template <int N>
using FileTensor = Tensor<N>;

#ifdef PARTENSOR_CONSTANTS_HPP
using Constraint = partensor::Constraint;
using Method = partensor::Method;
#else
// Constants.hpp:40-57, same values
enum class Method : uint8_t { als = 0, rnd = 1, bc = 2 };
enum class Constraint : uint8_t {
  unconstrained = 0, nonnegativity = 1, orthogonality = 2, sparsity = 3, constant = 4
};
#endif
using Duration = std::chrono::nanoseconds;  // PARTENSOR_basic.hpp:59

// Options (PARTENSOR_basic.hpp:283-336): the same fields, names and defaults
// (DefaultValues, :175-209).
template <typename Tensor_, typename ExecutionPolicy_ = execution::hip_policy>
struct Options {
  static constexpr std::size_t TnsSize = TensorTraits<Tensor_>::TnsSize;
  using MatrixType = typename TensorTraits<Tensor_>::MatrixType;
  using Constraints = std::array<Constraint, TnsSize>;
  using DoubleArray = std::array<double, TnsSize>;
  using StringArray = std::array<std::string, TnsSize>;

  Method method = Method::als;
  Constraints constraints;                       // nonnegativity
  double threshold_error = 1e-3;
  double nesterov_delta_1 = 1e-2;
  double nesterov_delta_2 = 1e-2;
  DoubleArray lambdas;                           // 0.1 (not consumed by the ALS path)
  unsigned max_iter = 20;
  Duration max_duration = Duration(10000);       // (not consumed by the ALS path)
  int accel_coeff = 3;
  int accel_fail = 0;
  bool acceleration = true;
  bool normalization = true;
  bool writeToFile = false;
  StringArray final_factors_paths;

  Options() {
    constraints.fill(Constraint::nonnegativity);
    lambdas.fill(0.1);
    for (std::size_t i = 0; i < TnsSize; ++i) final_factors_paths[i] = "final_" + std::to_string(i) + ".bin";
  }
};

// Status (PARTENSOR_basic.hpp:556-595).
template <typename Tensor_, typename ExecutionPolicy_ = execution::hip_policy>
struct Status {
  using MatrixArray = hip::MatrixArray<Tensor_>;
  Options<Tensor_, ExecutionPolicy_> options;
  double frob_tns = 0.0;
  double f_value = 0.0;
  double rel_costFunction = 0.0;
  unsigned ao_iter = 0;
  MatrixArray factors;
  Status() = default;
  explicit Status(const Options<Tensor_, ExecutionPolicy_> &opt) : options(opt) {}
};
#endif  // PARTENSOR_BASIC_HPP

}  // namespace hip

#ifdef PARTENSOR_BASIC_HPP
template <typename Tensor_>
using HipOptions = Options<Tensor_, execution::hip_policy, DefaultValues>;
namespace hip {
namespace detail {
template <typename Tensor_>
using StatusBase = partensor::Status<Tensor_, execution::hip_policy, DefaultValues>;
}  // namespace detail
}  // namespace hip
#else
template <typename Tensor_>
using HipOptions = hip::Options<Tensor_, execution::hip_policy>;
namespace hip {
namespace detail {
template <typename Tensor_>
using StatusBase = hip::Status<Tensor_, execution::hip_policy>;
}  // namespace detail
}  // namespace hip
#endif

// The policy's Status plus the device run's per-sweep log; converts to the Status type.
template <typename Tensor_>
struct HipStatus : hip::detail::StatusBase<Tensor_> {
  std::vector<ptn_sweep_log> sweeps;   // one record per outer iteration
  double sweep_seconds = 0.0;          // wall time of the outer loop (device-synchronised)
};

namespace hip {

// Owns one library handle (device memory, streams); not shared between threads.
class Handle {
 public:
  explicit Handle(int device = 0) {
    const ptn_status s = ptn_create(&h_, device);
    if (s != PTN_OK) throw std::runtime_error("ptn_create failed with status " + std::to_string((int)s));
  }
  ~Handle() { ptn_destroy(h_); }
  Handle(const Handle &) = delete;
  Handle &operator=(const Handle &) = delete;
  ptn_handle get() const { return h_; }
  void check(ptn_status s) const {
    if (s != PTN_OK) throw std::runtime_error(std::string("partensor hip: ") + ptn_last_error(h_));
  }

 private:
  ptn_handle h_ = nullptr;
};

namespace detail {

// Raw binary I/O in the reference's formats (ReadWrite.hpp:100-149, 307-341): a
// tensor is its column-major doubles, a factor matrix is written transposed (its
// row-major doubles).  Like the reference, failures are reported on stderr.
inline bool read_doubles(const std::string &path, size_t n, double *dst) {
  std::ifstream is(path, std::ios::binary);
  if (!is || !is.read(reinterpret_cast<char *>(dst), (std::streamsize)(n * sizeof(double)))) {
    std::cerr << "Exception opening/reading/closing file: " << path << std::endl;
    return false;
  }
  return true;
}

template <typename M>
void read_matrix_rowmajor(const std::string &path, M &A) {
  const int64_t rows = (int64_t)A.rows(), cols = (int64_t)A.cols();
  std::vector<double> t((size_t)(rows * cols));
  if (!read_doubles(path, t.size(), t.data())) return;
  for (int64_t i = 0; i < rows; ++i)
    for (int64_t j = 0; j < cols; ++j) A.data()[i + rows * j] = t[(size_t)(i * cols + j)];
}

template <typename M>
void write_matrix_rowmajor(const std::string &path, const M &A) {
  const int64_t rows = (int64_t)A.rows(), cols = (int64_t)A.cols();
  std::vector<double> t((size_t)(rows * cols));
  for (int64_t i = 0; i < rows; ++i)
    for (int64_t j = 0; j < cols; ++j) t[(size_t)(i * cols + j)] = A.data()[i + rows * j];
  std::ofstream os(path, std::ios::binary | std::ios::trunc);
  if (!os || !os.write(reinterpret_cast<const char *>(t.data()), (std::streamsize)(t.size() * sizeof(double))))
    std::cerr << "Exception opening/writing/closing file: " << path << std::endl;
}

// The options fields cpd() consumes, from either Options type, plus the policy's choices.
template <typename Opt>
ptn_options to_c(const Opt &o, const execution::hip_policy &pol, bool tree) {
  ptn_options c;
  ptn_default_options(&c);
  c.max_iter = (int)o.max_iter;
  c.threshold_error = o.threshold_error;
  c.nesterov_delta_1 = o.nesterov_delta_1;
  c.nesterov_delta_2 = o.nesterov_delta_2;
  c.accel_coeff = o.accel_coeff;
  c.accel_fail = o.accel_fail;
  c.acceleration = o.acceleration ? 1 : 0;
  c.normalization = o.normalization ? 1 : 0;
  for (std::size_t k = 0; k < o.constraints.size(); ++k) c.constraints[k] = (int)o.constraints[k];
  c.mttkrp_form = tree ? PTN_MTTKRP_DIMTREE : PTN_MTTKRP_STANDARD;
  c.arith = (int)pol.arith;
  c.max_inner = pol.max_inner;
  c.fuse_accel = pol.fuse_accel ? 1 : 0;
  return c;
}

// The ALS run behind every overload: X is uploaded once, factors start from `init`
// (column-major, caller's) or from the seeded U[0,1) generator.
template <typename Tensor_, typename InitArray>
HipStatus<Tensor_> run(const execution::hip_policy &pol, const int64_t *dims, const double *X, std::size_t R,
                       const HipOptions<Tensor_> &opt, const InitArray *init, bool tree) {
  constexpr std::size_t N = TensorTraits<Tensor_>::TnsSize;
  using MatrixType = typename TensorTraits<Tensor_>::MatrixType;
  Handle h(pol.device);
  const int storage = pol.arith == Arith::f64 ? PTN_TENSOR_F64 : PTN_TENSOR_F32;
  h.check(ptn_set_tensor(h.get(), (int)N, dims, X, storage));
  int64_t nf = 0;
  for (std::size_t k = 0; k < N; ++k) nf += dims[k] * (int64_t)R;
  std::vector<double> packed((size_t)nf), out((size_t)nf);
  if (init) {
    int64_t o = 0;
    for (std::size_t k = 0; k < N; ++k) {
      const auto &A = (*init)[k];
      if ((int64_t)A.rows() != dims[k] || (int64_t)A.cols() != (int64_t)R)
        throw std::runtime_error("partensor hip: initial factor " + std::to_string(k) + " has the wrong shape");
      for (int64_t e = 0; e < dims[k] * (int64_t)R; ++e) packed[(size_t)(o + e)] = A.data()[e];
      o += dims[k] * (int64_t)R;
    }
  } else {
    h.check(ptn_generate_factors(h.get(), (int)N, dims, (int)R, pol.init_seed, 0, packed.data()));
  }
  const ptn_options c = to_c(opt, pol, tree);
  HipStatus<Tensor_> st;
  st.options = opt;
  st.sweeps.resize(opt.max_iter);
  ptn_result res;
  h.check(ptn_cpd(h.get(), (int)R, packed.data(), &c, out.data(), &res, st.sweeps.data(), (int)st.sweeps.size()));
  st.sweeps.resize((size_t)res.ao_iter < st.sweeps.size() ? (size_t)res.ao_iter : st.sweeps.size());
  st.frob_tns = res.frob_sq;
  st.f_value = res.f_value;
  st.rel_costFunction = res.rel_cost;
  st.ao_iter = (unsigned)res.ao_iter;
  st.options.accel_coeff = res.accel_coeff;      // line_search_accel mutates these (Cpd.hpp:270-277)
  st.options.accel_fail = res.accel_fail;
  st.sweep_seconds = res.sweep_seconds;
  int64_t o = 0;
  for (std::size_t k = 0; k < N; ++k) {
    st.factors[k] = MatrixType(dims[k], (int64_t)R);
    for (int64_t e = 0; e < dims[k] * (int64_t)R; ++e) st.factors[k].data()[e] = out[(size_t)(o + e)];
    o += dims[k] * (int64_t)R;
  }
  if (opt.writeToFile)                           // writeFactorsToFile, Cpd.hpp:112-122
    for (std::size_t k = 0; k < N; ++k) write_matrix_rowmajor(opt.final_factors_paths[k], st.factors[k]);
  return st;
}

template <typename Tensor_>
std::array<int64_t, TensorTraits<Tensor_>::TnsSize> dims_of(const Tensor_ &X) {
  std::array<int64_t, TensorTraits<Tensor_>::TnsSize> d;
  for (std::size_t k = 0; k < d.size(); ++k) d[k] = (int64_t)X.dimension((int)k);
  return d;
}

// A tensor of the given dimensions (this header's, or Eigen's through resize(array)).
template <typename T, std::size_t N>
T make_tensor(const std::array<int64_t, N> &d) {
  T X;
  std::array<typename std::decay<decltype(X.dimension(0))>::type, N> dd;
  for (std::size_t k = 0; k < N; ++k) dd[k] = d[k];
  X.resize(dd);
  return X;
}

template <std::size_t N>
FileTensor<(int)N> read_tensor(const std::array<int, N> &dims, const std::string &path) {
  std::array<int64_t, N> d;
  int64_t n = 1;
  for (std::size_t k = 0; k < N; ++k) { d[k] = dims[k]; n *= dims[k]; }
  auto X = make_tensor<FileTensor<(int)N>>(d);
  read_doubles(path, (size_t)n, X.data());
  return X;
}

template <std::size_t N>
MatrixArray<FileTensor<(int)N>> read_factor_files(const std::array<int, N> &dims, std::size_t R,
                                                  const std::string *paths) {
  using MatrixType = typename TensorTraits<FileTensor<(int)N>>::MatrixType;
  MatrixArray<FileTensor<(int)N>> A;
  for (std::size_t k = 0; k < N; ++k) {
    A[k] = MatrixType((int64_t)dims[k], (int64_t)R);
    read_matrix_rowmajor(paths[k], A[k]);
  }
  return A;
}

// X = [[A]] from factor files (the true_paths form: CpdGen of the true factors)
template <std::size_t N>
FileTensor<(int)N> tensor_from_factor_files(const std::array<int, N> &dims, std::size_t R,
                                            const std::array<std::string, N> &paths) {
  const auto A = read_factor_files<N>(dims, R, paths.data());
  std::array<int64_t, N> d;
  int64_t n = 1;
  for (std::size_t k = 0; k < N; ++k) { d[k] = dims[k]; n *= dims[k]; }
  auto X = make_tensor<FileTensor<(int)N>>(d);
  std::vector<int64_t> idx(N, 0);
  for (int64_t e = 0; e < n; ++e) {
    double s = 0.0;
    for (std::size_t r = 0; r < R; ++r) {
      double p = 1.0;
      for (std::size_t k = 0; k < N; ++k) p *= A[k].data()[idx[k] + d[k] * (int64_t)r];
      s += p;
    }
    X.data()[e] = s;
    for (std::size_t k = 0; k < N && ++idx[k] == d[k]; ++k) idx[k] = 0;
  }
  return X;
}

template <typename Tensor_>
HipStatus<Tensor_> dispatch(const execution::hip_policy &pol, const Tensor_ &X, std::size_t R,
                            const HipOptions<Tensor_> &opt, const MatrixArray<Tensor_> *init, bool tree) {
  const auto d = dims_of(X);
  return run<Tensor_>(pol, d.data(), X.data(), R, opt, init, tree);
}

template <std::size_t N>
using FileStatus = HipStatus<FileTensor<(int)N>>;
template <std::size_t N>
using FileOptions = HipOptions<FileTensor<(int)N>>;

}  // namespace detail
}  // namespace hip

// ---- cpd(execution::hip, ...) : Cpd.hpp overload set --------------------------------
template <typename Tensor_>
HipStatus<Tensor_> cpd(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, HipOptions<Tensor_>(), nullptr, false);
}
template <typename Tensor_>
HipStatus<Tensor_> cpd(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R,
                       const HipOptions<Tensor_> &options) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, options, nullptr, false);
}
template <typename Tensor_>
HipStatus<Tensor_> cpd(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R,
                       const hip::MatrixArray<Tensor_> &factorsInit) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, HipOptions<Tensor_>(), &factorsInit, false);
}
template <typename Tensor_>
HipStatus<Tensor_> cpd(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R,
                       const HipOptions<Tensor_> &options, const hip::MatrixArray<Tensor_> &factorsInit) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, options, &factorsInit, false);
}

// file forms: tensor file; tensor + initial factor files; true + initial factor files
template <std::size_t N>
hip::detail::FileStatus<N> cpd(execution::hip_policy pol, const std::array<int, N> &tnsDims, std::size_t R,
                               const std::string &path,
                               const hip::detail::FileOptions<N> &options = hip::detail::FileOptions<N>()) {
  const auto X = hip::detail::read_tensor(tnsDims, path);
  return hip::detail::dispatch(pol, X, R, options, nullptr, false);
}
template <std::size_t N>
hip::detail::FileStatus<N> cpd(execution::hip_policy pol, const std::array<int, N> &tnsDims, std::size_t R,
                               const std::array<std::string, N + 1> &paths,
                               const hip::detail::FileOptions<N> &options = hip::detail::FileOptions<N>()) {
  const auto X = hip::detail::read_tensor(tnsDims, paths[0]);
  const auto init = hip::detail::read_factor_files<N>(tnsDims, R, paths.data() + 1);
  return hip::detail::dispatch(pol, X, R, options, &init, false);
}
template <std::size_t N>
hip::detail::FileStatus<N> cpd(execution::hip_policy pol, const std::array<int, N> &tnsDims, std::size_t R,
                               const std::array<std::string, N> &true_paths,
                               const std::array<std::string, N> &init_paths,
                               const hip::detail::FileOptions<N> &options) {
  const auto X = hip::detail::tensor_from_factor_files<N>(tnsDims, R, true_paths);
  const auto init = hip::detail::read_factor_files<N>(tnsDims, R, init_paths.data());
  return hip::detail::dispatch(pol, X, R, options, &init, false);
}

// ---- cpdDimTree(execution::hip, ...) : CpdDimTree.hpp overload set (orders 3 to 8) ----
template <typename Tensor_>
HipStatus<Tensor_> cpdDimTree(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, HipOptions<Tensor_>(), nullptr, true);
}
template <typename Tensor_>
HipStatus<Tensor_> cpdDimTree(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R,
                              const HipOptions<Tensor_> &options) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, options, nullptr, true);
}
template <typename Tensor_>
HipStatus<Tensor_> cpdDimTree(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R,
                              const hip::MatrixArray<Tensor_> &factorsInit) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, HipOptions<Tensor_>(), &factorsInit, true);
}
template <typename Tensor_>
HipStatus<Tensor_> cpdDimTree(execution::hip_policy pol, const Tensor_ &tnsX, std::size_t R,
                              const HipOptions<Tensor_> &options, const hip::MatrixArray<Tensor_> &factorsInit) {
  return hip::detail::dispatch<Tensor_>(pol, tnsX, R, options, &factorsInit, true);
}
template <std::size_t N>
hip::detail::FileStatus<N> cpdDimTree(execution::hip_policy pol, const std::array<int, N> &tnsDims, std::size_t R,
                                      const std::string &path,
                                      const hip::detail::FileOptions<N> &options = hip::detail::FileOptions<N>()) {
  const auto X = hip::detail::read_tensor(tnsDims, path);
  return hip::detail::dispatch(pol, X, R, options, nullptr, true);
}
template <std::size_t N>
hip::detail::FileStatus<N> cpdDimTree(execution::hip_policy pol, const std::array<int, N> &tnsDims, std::size_t R,
                                      const std::array<std::string, N + 1> &paths,
                                      const hip::detail::FileOptions<N> &options = hip::detail::FileOptions<N>()) {
  const auto X = hip::detail::read_tensor(tnsDims, paths[0]);
  const auto init = hip::detail::read_factor_files<N>(tnsDims, R, paths.data() + 1);
  return hip::detail::dispatch(pol, X, R, options, &init, true);
}
template <std::size_t N>
hip::detail::FileStatus<N> cpdDimTree(execution::hip_policy pol, const std::array<int, N> &tnsDims, std::size_t R,
                                      const std::array<std::string, N> &true_paths,
                                      const std::array<std::string, N> &init_paths,
                                      const hip::detail::FileOptions<N> &options) {
  const auto X = hip::detail::tensor_from_factor_files<N>(tnsDims, R, true_paths);
  const auto init = hip::detail::read_factor_files<N>(tnsDims, R, init_paths.data());
  return hip::detail::dispatch(pol, X, R, options, &init, true);
}

}  // namespace partensor

#endif  // PARTENSOR_HIP_HPP
