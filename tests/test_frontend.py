"""C++ front end (execution::hip policy over the C ABI).

CPU: the header compiles on its own and every overload of the cpd / cpdDimTree
set resolves (a translation unit instantiating each one, compiled with g++); with
the reference's own execution.hpp and Constants.hpp included first, the policy is an
execution policy of the reference's trait, options set as reference user code sets
them (method, constraints, lambdas, max_duration) compile, and cpd(execution::hip, ...)
resolves to the hip overloads and links against the library.
GPU: the example program runs the file-based overloads on the C1 tensor and
initial factors written in the reference's formats (tensor column-major, factors
row-major), and its Status and written factors equal the Python binding's run of
the same decomposition bit for bit, and the FP64 oracle's to 1e-12."""
import json
import os
import subprocess
import textwrap

import numpy as np
import pytest

import oracle as o
import ptn_hip as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FRONT = os.path.join(ROOT, "partensor-toolbox_amd", "frontend")
INC = os.path.join(ROOT, "include")

OVERLOADS = textwrap.dedent("""
    #include "partensor_hip.hpp"
    using T3 = partensor::hip::Tensor<3>;
    using Opt = partensor::HipOptions<T3>;
    void instantiate() {
      namespace ex = partensor::execution;
      T3 X({2, 3, 4});
      Opt opt;
      partensor::hip::MatrixArray<T3> init;
      std::array<int, 3> d{2, 3, 4};
      std::array<std::string, 4> p4;
      std::array<std::string, 3> p3;
      static_assert(decltype(ex::hip_policy::__allow_hip())::value, "hip tag");
      constexpr auto fast = ex::hip.with_arith(partensor::hip::Arith::bf16x3).with_fused_line_search();
      static_assert(fast.fuse_accel && !ex::hip.fuse_accel, "policy choices are per object");
      (void)partensor::cpd(ex::hip, X, 2);
      (void)partensor::cpdDimTree(fast, X, 2, opt);
      partensor::hip::Status<T3> base = partensor::cpd(ex::hip.on_device(0), X, 2, opt);
      (void)base;
      (void)partensor::cpd(ex::hip, X, 2, opt);
      (void)partensor::cpd(ex::hip, X, 2, init);
      (void)partensor::cpd(ex::hip, X, 2, opt, init);
      (void)partensor::cpd(ex::hip, d, 2, std::string("x.bin"));
      (void)partensor::cpd(ex::hip, d, 2, std::string("x.bin"), opt);
      (void)partensor::cpd(ex::hip, d, 2, p4);
      (void)partensor::cpd(ex::hip, d, 2, p4, opt);
      (void)partensor::cpd(ex::hip, d, 2, p3, p3, opt);
      (void)partensor::cpdDimTree(ex::hip, X, 2);
      (void)partensor::cpdDimTree(ex::hip, X, 2, opt);
      (void)partensor::cpdDimTree(ex::hip, X, 2, init);
      (void)partensor::cpdDimTree(ex::hip, X, 2, opt, init);
      (void)partensor::cpdDimTree(ex::hip, d, 2, std::string("x.bin"), opt);
      (void)partensor::cpdDimTree(ex::hip, d, 2, p4, opt);
      (void)partensor::cpdDimTree(ex::hip, d, 2, p3, p3, opt);
    }
""")


def test_front_end_overload_set_compiles(tmp_path):
    src = tmp_path / "overloads.cpp"
    src.write_text(OVERLOADS)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", INC, "-I", FRONT,
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


REF_INCLUDE = "/root/reference/include"

WITH_REFERENCE_HEADERS = textwrap.dedent("""
    #include <cstdint>
    #include "execution.hpp"        // the reference's policy tags and trait
    #include "Constants.hpp"        // the reference's Constraint / Method enums
    #include "partensor_hip.hpp"
    #include <chrono>
    namespace ex = partensor::execution;
    static_assert(ex::is_execution_policy_v<ex::hip_policy>, "hip is an execution policy");
    static_assert(std::is_same<ex::internal::enable_if_execution_policy<const ex::hip_policy &, int>, int>::value,
                  "the reference's SFINAE gate accepts the hip tag");
    static_assert(!ex::is_execution_policy_v<int>, "");
    using T3 = partensor::hip::Tensor<3>;
    using Options = partensor::HipOptions<T3>;
    using Constraint = partensor::Constraint;
    static_assert(std::is_same<Options::Constraints, std::array<Constraint, 3>>::value, "reference enum");
    int main() {
      T3 X({4, 5, 6});
      Options options;                                   // user code as test/cpd_options.cpp writes it
      options.method = partensor::Method::als;
      options.max_iter = 50;
      options.constraints[0] = Constraint::nonnegativity;
      options.constraints[1] = Constraint::unconstrained;
      options.constraints[2] = Constraint::orthogonality;
      options.lambdas[0] = 0.5;
      options.max_duration = std::chrono::nanoseconds(5000);
      partensor::HipStatus<T3> st = partensor::cpd(ex::hip, X, 2, options);
      auto st2 = partensor::cpdDimTree(ex::hip, X, 2, options);
      std::array<int, 3> dims{4, 5, 6};
      auto st3 = partensor::cpd(ex::hip, dims, 2, std::string("x.bin"), options);
      return (int)(st.ao_iter + st2.ao_iter + st3.ao_iter);
    }
""")


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_INCLUDE, "execution.hpp")),
                    reason="the reference's headers are not on this machine")
def test_front_end_with_reference_headers(tmp_path):
    """The drop-in inside a reference build: the reference's execution.hpp (policy tags,
    is_execution_policy, enable_if_execution_policy; execution.hpp:170-327) and
    Constants.hpp (Constraint, Method) come first; the front end specialises the trait
    and compiles next to them, and the program links against libptn_hip.so."""
    src = tmp_path / "with_reference.cpp"
    src.write_text(WITH_REFERENCE_HEADERS)
    exe = tmp_path / "with_reference"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", REF_INCLUDE, "-I", INC, "-I", FRONT,
                        str(src), "-o", str(exe), "-L", os.path.join(ROOT, "partensor-toolbox_amd"), "-lptn_hip"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _write_inputs(tmp_path, X, init):
    paths = [str(tmp_path / "tensor.bin")] + [str(tmp_path / ("init_%d.bin" % k)) for k in range(len(init))]
    np.asfortranarray(X).ravel(order="F").tofile(paths[0])             # column-major doubles
    for k, A in enumerate(init):
        np.ascontiguousarray(A).ravel(order="C").tofile(paths[k + 1])   # factors row-major
    return paths


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["cpd", "dimtree", "dimtree-bf16x3"])
def test_front_end_example_matches_binding_and_oracle(tmp_path, form):
    exe = os.path.join(FRONT, "cpd_example")
    assert os.path.exists(exe), "build the package first (make in partensor-toolbox_amd)"
    dims, R, sweeps = [100, 100, 100], 10, 20
    X, _ = o.make_tensor(dims, R, seed=0, snr_db=30.0)
    init = o.gen_factors(dims, R, seed=1)
    paths = _write_inputs(tmp_path, X, init)
    prefix = str(tmp_path / "final_")
    print("[gpu-test] front end", form, flush=True)
    r = subprocess.run([exe, form, "3", str(R), str(sweeps)] + [str(d) for d in dims] + paths + [prefix],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    st = json.loads(r.stdout.strip().splitlines()[-1])
    F = [np.fromfile(prefix + "%d.bin" % k).reshape(dims[k], R) for k in range(3)]   # row-major on disk
    kw = dict(max_iter=sweeps, threshold_error=0.0)
    with P.Handle(0) as h:
        if form == "dimtree-bf16x3":
            h.set_tensor(X, storage=P.TENSOR_F32)
            Fb, stb, logb = h.cpd(R, init, mttkrp_form=P.MTTKRP_DIMTREE, arith=P.ARITH_BF16X3, fuse_accel=1, **kw)
        else:
            h.set_tensor(X)
            Fb, stb, logb = (h.cpd_dimtree if form == "dimtree" else h.cpd)(R, init, **kw)
    assert st["ao_iter"] == stb["ao_iter"] == sweeps
    assert st["f_value"] == stb["f_value"]
    assert (st["accel_coeff"], st["accel_fail"]) == (stb["accel_coeff"], stb["accel_fail"])
    for a, b in zip(F, Fb):
        assert np.array_equal(a, b)
    if form == "dimtree-bf16x3":
        return                                 # the FP32-stream forms are checked against the oracle elsewhere
    force = [l["accel_accepted"] if l["accel_tried"] else -1 for l in logb]
    Fo, sto, logo = o.cpd(X, R, init, cost_form=o.COST_GRAM, force_accept=force,
                          mttkrp_form=o.MTTKRP_DIMTREE if form == "dimtree" else o.MTTKRP_STANDARD, **kw)
    for l in logo:                             # a forced outcome must be a rounding-level tie
        assert not l["accel_overridden"] or abs(l["f_value"] - l["f_accel"]) <= 1e-12 * l["f_value"]
    assert abs(st["f_value"] ** 2 - sto["f_value"] ** 2) < 2e-13 * sto["frob_sq"]
    for a, b in zip(F, Fo):
        assert P.rel_err(a, b) < 1e-12
