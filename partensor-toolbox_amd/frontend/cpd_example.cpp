// cpd_example: runs partensor::cpd / cpdDimTree with the hip policy on a tensor and
// initial factors read from files in the reference's formats (the (dims, R, paths,
// options) overload, Cpd.hpp:1489-1575), writes the final factors with writeToFile
// and prints the Status as one JSON line.  "dimtree-bf16x3" runs the tree with the
// FP32 stream, bf16x3 contractions and the fused line-search pass (the policy object's
// choices).
//
//   cpd_example cpd|dimtree|dimtree-bf16x3 N R max_iter I_0 .. I_{N-1} tensor.bin
//               init_0.bin .. init_{N-1}.bin out_prefix
#include "partensor_hip.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>

template <std::size_t N>
int run(const std::string &form, std::size_t R, unsigned max_iter, char **argv) {
  std::array<int, N> dims;
  for (std::size_t k = 0; k < N; ++k) dims[k] = std::atoi(argv[k]);
  std::array<std::string, N + 1> paths;
  for (std::size_t k = 0; k < N + 1; ++k) paths[k] = argv[N + k];
  const std::string prefix = argv[2 * N + 1];
  partensor::HipOptions<partensor::hip::Tensor<(int)N>> opt;
  opt.max_iter = max_iter;
  opt.threshold_error = 0.0;
  opt.writeToFile = true;
  for (std::size_t k = 0; k < N; ++k) opt.final_factors_paths[k] = prefix + std::to_string(k) + ".bin";
  namespace ex = partensor::execution;
  const auto st = form == "cpd"       ? partensor::cpd(ex::hip, dims, R, paths, opt)
                  : form == "dimtree" ? partensor::cpdDimTree(ex::hip, dims, R, paths, opt)
                                      : partensor::cpdDimTree(ex::hip.with_arith(partensor::hip::Arith::bf16x3)
                                                                  .with_fused_line_search(),
                                                              dims, R, paths, opt);
  std::printf("{\"ao_iter\": %u, \"f_value\": %.17g, \"rel_costFunction\": %.17g, \"frob_tns\": %.17g, "
              "\"accel_coeff\": %d, \"accel_fail\": %d, \"nnls_iters_first\": %d}\n",
              st.ao_iter, st.f_value, st.rel_costFunction, st.frob_tns, st.options.accel_coeff,
              st.options.accel_fail, st.sweeps.empty() ? -1 : st.sweeps[0].nnls_iters[0]);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s cpd|dimtree|dimtree-bf16x3 N R max_iter dims... tensor init... out_prefix\n",
                 argv[0]);
    return 2;
  }
  const std::string form = argv[1];
  if (form != "cpd" && form != "dimtree" && form != "dimtree-bf16x3") {
    std::fprintf(stderr, "unknown form %s\n", argv[1]);
    return 2;
  }
  const int N = std::atoi(argv[2]);
  const std::size_t R = (std::size_t)std::atoi(argv[3]);
  const unsigned max_iter = (unsigned)std::atoi(argv[4]);
  if (argc != 5 + 2 * N + 2) {
    std::fprintf(stderr, "expected %d arguments after max_iter\n", 2 * N + 2);
    return 2;
  }
  try {
    if (N == 3) return run<3>(form, R, max_iter, argv + 5);
    if (N == 4) return run<4>(form, R, max_iter, argv + 5);
    std::fprintf(stderr, "this example handles orders 3 and 4\n");
    return 2;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
