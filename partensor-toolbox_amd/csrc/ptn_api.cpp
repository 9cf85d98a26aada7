// C ABI of the MI355X execution policy: handle, tensor upload / generation, the
// ALS driver (cpd() / cpdDimTree() semantics) and single-operation entry points.
// Host control flow follows CPD::als (Cpd.hpp:288-356), CPD<T,cuda>::als
// (CpdCUDA.hpp:277-383) and CPD_DIMTREE::als (CpdDimTree.hpp:317-401); every
// tensor-sized and R x R computation runs in the HIP kernels.
#include "../../include/ptn_hip.h"
#include "ptn_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <vector>

using namespace ptn;

// The tensors of a batch of small problems (ptn_batch_set / ptn_batch_generate): FP32,
// back to back, each starting on a 64-element boundary.
struct BatchSet {
  int nprob = 0;
  std::vector<BatchDesc> desc;   // host copy
  int64_t nx = 0, nrows = 0;     // tensor elements (padded), factor rows over all problems
  float *x32 = nullptr;
  ~BatchSet() {
    if (x32) (void)hipFree(x32);
  }
};

namespace {
struct MlpState;
}

struct ptn_handle_s {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;                // eigen / momentum set-up of the next NNLS, beside the contractions
  hipStream_t aux = nullptr;                 // the order-3 tree's right node, beside mode 1
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // [0] main -> side, [1] side -> main,
                                                             // [2] main -> aux, [3] aux -> main
  std::string err;
  int N = 0;
  int64_t dims[kMaxOrder] = {0};
  int64_t total = 0;
  double *x64 = nullptr;
  float *x32 = nullptr;
  double *scal = nullptr;    // device scalars: [0] ||X||^2, [1..] scratch
  double *part = nullptr;    // reduction partials for norms
  bool frob_valid = false;
  double frob_host = 0.0;
  BatchSet *batch = nullptr;
  MlpState *mlp = nullptr;    // the warm-start network, its training set and work buffers
};

// Device scratch of ptn_nesterov_mnls_noalloc: a file-scope device array, one copy per GPU,
// held until the process ends.  Host code never uses the symbol itself as an address: the
// device address of each GPU's copy is looked up once per device (ptn_create, with the
// device current) and kept in noalloc_ptr[device].
constexpr int kNoallocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[PTN_NOALLOC_SCRATCH_BYTES];
static std::mutex noalloc_mu;
static void *noalloc_ptr[kNoallocMaxDevices] = {};

namespace {

constexpr int kNormParts = 1024;
constexpr int64_t kTensorPad = 64;

struct Fail {
  ptn_status code;
};

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      h->err = std::string(#expr) + ": " + hipGetErrorString(e_);                        \
      throw Fail{e_ == hipErrorOutOfMemory ? PTN_ERR_OOM : PTN_ERR_HIP};                  \
    }                                                                                     \
  } while (0)

#define REQUIRE(cond, code, msg)        \
  do {                                  \
    if (!(cond)) {                      \
      h->err = (msg);                   \
      throw Fail{code};                 \
    }                                   \
  } while (0)

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  ptn_handle h = nullptr;
  DevBuf() = default;
  DevBuf(ptn_handle hh, size_t count) { alloc(hh, count); }
  void alloc(ptn_handle hh, size_t count) {
    h = hh;
    n = count ? count : 1;
    HIPCHK(hipMalloc((void **)&p, n * sizeof(T)));
    HIPCHK(hipMemsetAsync(p, 0, n * sizeof(T), h->stream));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  operator T *() const { return p; }
};

template <typename T>
struct PinnedBuf {      // page-locked, coherent host memory that kernels write directly (dev = its device address)
  T *p = nullptr;
  T *dev = nullptr;
  explicit PinnedBuf(ptn_handle h) {
    HIPCHK(hipHostMalloc((void **)&p, sizeof(T), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset((void *)p, 0, sizeof(T));
    HIPCHK(hipHostGetDevicePointer((void **)&dev, p, 0));
  }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
};

struct EventPair {                           // two timing-free events, destroyed with the scope
  hipEvent_t e[2] = {nullptr, nullptr};
  explicit EventPair(ptn_handle h) {
    for (hipEvent_t &x : e) HIPCHK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  }
  ~EventPair() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
  EventPair(const EventPair &) = delete;
  EventPair &operator=(const EventPair &) = delete;
};

ptn_status run(ptn_handle h, const std::function<void()> &fn);

// ------------------------------------------------------------------------------------
// the warm-start network: parameters, Adam state, training set, activations
// ------------------------------------------------------------------------------------
uint16_t f32_to_bf16(float f) {               // round to nearest even; NaN stays NaN
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct MlpState {
  ptn_handle h;
  bool ready = false;
  DevBuf<float> w, g, m1, m2;                 // FP32 master parameters, gradient, Adam moments
  DevBuf<uint16_t> wb, w2t, w3t;              // bf16 W1 | W2 | W3, and W2^T, W3^T
  int step = 0;
  DevBuf<uint16_t> dx;                        // training set: features (bf16), targets, ranks
  DevBuf<float> dt;
  DevBuf<unsigned char> drank;
  std::vector<unsigned char> hrank;
  int64_t n = 0, cap = 0;
  int rows = 0;                               // activation buffers hold this many rows
  DevBuf<uint16_t> X, XT, H1, H2, H1T, H2T, dYb, dYT, dH1, dH1T, dH2, dH2T;
  DevBuf<float> Y, T, dYf, scale;
  DevBuf<unsigned char> rank;
  DevBuf<double> start, part;
  explicit MlpState(ptn_handle hh) : h(hh) {
    w.alloc(h, kMlpParams); g.alloc(h, kMlpParams); m1.alloc(h, kMlpParams); m2.alloc(h, kMlpParams);
    wb.alloc(h, kMlpWb); w2t.alloc(h, (size_t)kMlpH * kMlpH); w3t.alloc(h, (size_t)kMlpH * kMlpR);
    part.alloc(h, 64);
  }
  template <typename T_>
  static void grow(ptn_handle h, DevBuf<T_> &b, size_t count) {
    if (b.p && b.n >= count) return;
    b.~DevBuf();
    new (&b) DevBuf<T_>();
    b.alloc(h, count);
  }
  void ensure_rows(int r) {
    r = (r + 31) / 32 * 32;
    if (r <= rows) return;
    HIPCHK(hipDeviceSynchronize());
    const size_t rin = (size_t)r * kMlpIn, rh = (size_t)r * kMlpH, ro = (size_t)r * kMlpR;
    grow(h, X, rin); grow(h, XT, rin); grow(h, H1, rh); grow(h, H2, rh); grow(h, H1T, rh); grow(h, H2T, rh);
    grow(h, dYb, ro); grow(h, dYT, ro); grow(h, dH1, rh); grow(h, dH1T, rh); grow(h, dH2, rh); grow(h, dH2T, rh);
    grow(h, Y, ro); grow(h, T, ro); grow(h, dYf, ro); grow(h, scale, (size_t)r); grow(h, rank, (size_t)r);
    grow(h, start, ro);
    rows = r;
  }
  void ensure_data(int64_t need) {
    if (need <= cap) return;
    int64_t nc = std::max<int64_t>(need, 2 * cap);
    HIPCHK(hipDeviceSynchronize());
    DevBuf<uint16_t> nx(h, (size_t)nc * kMlpIn);
    DevBuf<float> nt(h, (size_t)nc * kMlpR);
    DevBuf<unsigned char> nr(h, (size_t)nc);
    if (n > 0) {
      HIPCHK(hipMemcpy(nx.p, dx.p, (size_t)n * kMlpIn * 2, hipMemcpyDeviceToDevice));
      HIPCHK(hipMemcpy(nt.p, dt.p, (size_t)n * kMlpR * 4, hipMemcpyDeviceToDevice));
      HIPCHK(hipMemcpy(nr.p, drank.p, (size_t)n, hipMemcpyDeviceToDevice));
    }
    std::swap(dx.p, nx.p); std::swap(dx.n, nx.n); dx.h = h;
    std::swap(dt.p, nt.p); std::swap(dt.n, nt.n); dt.h = h;
    std::swap(drank.p, nr.p); std::swap(drank.n, nr.n); drank.h = h;
    cap = nc;
  }
  // Y = network(X) for B rows (activations kept for the backward pass)
  void forward(hipStream_t s, int B) {
    gemm_nt(s, B, kMlpH, kMlpIn, X.p, kMlpIn, wb.p + kMlpW1b, kMlpIn, w.p + kMlpB1, nullptr, 0, 1, nullptr, 0, H1.p,
            kMlpH);
    gemm_nt(s, B, kMlpH, kMlpH, H1.p, kMlpH, wb.p + kMlpW2b, kMlpH, w.p + kMlpB2, nullptr, 0, 1, nullptr, 0, H2.p,
            kMlpH);
    gemm_nt(s, B, kMlpR, kMlpH, H2.p, kMlpH, wb.p + kMlpW3b, kMlpH, w.p + kMlpB3, nullptr, 0, 0, Y.p, kMlpR, nullptr,
            0);
  }
  // gradient of the masked mean squared error into g (B a multiple of 32); per-block
  // partial losses in part[0..64)
  void backward(hipStream_t s, int B, float inv_count) {
    mse_grad(s, B, Y.p, T.p, rank.p, inv_count, dYb.p, dYf.p, part.p, 64);
    colsum_f32(s, B, kMlpR, dYf.p, kMlpR, g.p + kMlpB3);
    transpose_bf16(s, B, kMlpR, dYb.p, kMlpR, dYT.p, B);
    transpose_bf16(s, B, kMlpH, H2.p, kMlpH, H2T.p, B);
    gemm_nt(s, kMlpR, kMlpH, B, dYT.p, B, H2T.p, B, nullptr, nullptr, 0, 0, g.p + kMlpW3, kMlpH, nullptr, 0);
    transpose_bf16(s, kMlpR, kMlpH, wb.p + kMlpW3b, kMlpH, w3t.p, kMlpR);
    gemm_nt(s, B, kMlpH, kMlpR, dYb.p, kMlpR, w3t.p, kMlpR, nullptr, H2.p, kMlpH, 0, nullptr, 0, dH2.p, kMlpH);
    colsum_bf16(s, B, kMlpH, dH2.p, kMlpH, g.p + kMlpB2);
    transpose_bf16(s, B, kMlpH, dH2.p, kMlpH, dH2T.p, B);
    transpose_bf16(s, B, kMlpH, H1.p, kMlpH, H1T.p, B);
    gemm_nt(s, kMlpH, kMlpH, B, dH2T.p, B, H1T.p, B, nullptr, nullptr, 0, 0, g.p + kMlpW2, kMlpH, nullptr, 0);
    transpose_bf16(s, kMlpH, kMlpH, wb.p + kMlpW2b, kMlpH, w2t.p, kMlpH);
    gemm_nt(s, B, kMlpH, kMlpH, dH2.p, kMlpH, w2t.p, kMlpH, nullptr, H1.p, kMlpH, 0, nullptr, 0, dH1.p, kMlpH);
    colsum_bf16(s, B, kMlpH, dH1.p, kMlpH, g.p + kMlpB1);
    transpose_bf16(s, B, kMlpH, dH1.p, kMlpH, dH1T.p, B);
    transpose_bf16(s, B, kMlpIn, X.p, kMlpIn, XT.p, B);
    gemm_nt(s, kMlpH, kMlpIn, B, dH1T.p, B, XT.p, B, nullptr, nullptr, 0, 0, g.p + kMlpW1, kMlpIn, nullptr, 0);
  }
};

MlpState &mlp_of(ptn_handle h) {
  if (!h->mlp) h->mlp = new MlpState(h);
  return *h->mlp;
}

int round_up16(int R) { return (R + 15) / 16 * 16; }

int64_t prod(const int64_t *d, int n) {
  int64_t p = 1;
  for (int k = 0; k < n; ++k) p *= d[k];
  return p;
}

void check_dims(ptn_handle h, int N, const int64_t *dims) {
  REQUIRE(N >= 3 && N <= kMaxOrder, PTN_ERR_INVALID, "tensor order must be in [3, 8]");
  REQUIRE(dims != nullptr, PTN_ERR_INVALID, "dims is null");
  for (int k = 0; k < N; ++k) REQUIRE(dims[k] >= 1 && dims[k] < (1ll << 31), PTN_ERR_INVALID, "bad dimension");
  REQUIRE(prod(dims, N) / dims[0] < (1ll << 32), PTN_ERR_UNSUPPORTED, "prod of dims 1..N-1 must be < 2^32");
}

// ------------------------------------------------------------------------------------
// working set of one decomposition
// ------------------------------------------------------------------------------------
struct Work {
  ptn_handle h;
  int N, R, ldr, rpf;
  int64_t dims[kMaxOrder];
  int64_t total, maxI, fsz;
  int64_t off[kMaxOrder];
  int arith;
  bool f32;                                  // FP32 tensor stream (arith F32 / BF16X3 / F32_F64)
  bool mfma32;                               // contractions on FP32-output kernels (F32 / BF16X3)
  DevBuf<double> fac, old, acc, grams, agr, Vm, Vacc, M, M2, Aold, colsum, part, T, YL, nscr, scal, hist, betas,
      gpart, gpart_side, cscr;
  DevBuf<float> T32, YL32, YL2, part32, wf;
  DevBuf<uint16_t> whi, wlo;
  DevBuf<NnlsParams> prm;
  DevBuf<uint64_t> mask;
  DevBuf<int> flags;                         // [0] line-search step accepted, [1] threshold stop
  int64_t part_cap = 0, part32_cap = 0, wcap = 0, hist_cap = 0;
  struct NodeRef {
    const double *p64 = nullptr;
    const float *p32 = nullptr;
    int64_t ld = 0;
    const int *sel = nullptr;
    int64_t sel_off = 0;
  } left, right;                             // dimension-tree nodes cached between the modes of a sweep
  bool left_ready = false;                   // left node already formed for the coming mode 0 (line-search pass)
  bool right_ready = false;                  // order-3 right node X x_0 A_0 already formed (aux stream)
  bool m0_ready = false;                     // mode-0 MTTKRP already in M (paired pass, chosen on the device)
  // time_passes: events around the sweep's tensor passes, slot = sweep parity ([2 q] start,
  // [2 q + 1] end of pass q: 0 left node, 1 right node); tslot < 0 when not timing
  hipEvent_t (*kev)[4] = nullptr;
  bool kev_used[2][2] = {{false, false}, {false, false}};
  int tslot = -1;
  void pass_mark(int q, bool end, hipStream_t s) {
    if (tslot < 0 || q < 0) return;
    HIPCHK(hipEventRecord(kev[tslot][2 * q + (end ? 1 : 0)], s));
    if (end) kev_used[tslot][q] = true;
  }

  Work(ptn_handle hh, int R_, int arith_, bool need_tree)
      : h(hh), N(hh->N), R(R_), ldr(round_up16(R_)), rpf((R_ + 31) / 32 * 32), arith(arith_) {
    f32 = arith != PTN_ARITH_F64;
    mfma32 = arith == PTN_ARITH_F32 || arith == PTN_ARITH_BF16X3;
    for (int k = 0; k < N; ++k) dims[k] = h->dims[k];
    total = h->total;
    maxI = 0;
    int64_t o = 0;
    for (int k = 0; k < N; ++k) {
      off[k] = o;
      o += dims[k] * ldr;
      if (dims[k] > maxI) maxI = dims[k];
    }
    fsz = o;
    fac.alloc(h, fsz); old.alloc(h, fsz); acc.alloc(h, fsz);
    grams.alloc(h, (size_t)N * R * R); agr.alloc(h, (size_t)N * R * R);
    Vm.alloc(h, (size_t)(N + 1) * R * R); Vacc.alloc(h, (size_t)R * R);
    M.alloc(h, maxI * ldr); Aold.alloc(h, maxI * ldr); colsum.alloc(h, kMaxRank);
    prm.alloc(h, kMaxOrder + 1); mask.alloc(h, nnls_mask_words(maxI)); flags.alloc(h, 4);
    nscr.alloc(h, 2 * kMaxOrder * kMaxRank); scal.alloc(h, 8); betas.alloc(h, (size_t)(N + 1) * (kMaxInner + 1));
    gpart.alloc(h, (size_t)((maxI + 31) / 32) * R * R); gpart_side.alloc(h, (size_t)((maxI + 31) / 32) * R * R);
    cscr.alloc(h, (size_t)3 * R * R + R);    // R x R work of the unconstrained / orthogonality updates
    const int64_t C = total / dims[0];
    const int64_t yl = left_rows();
    // reduction partials: mode-0 split-K (<= 256 splits of I_0 rows) and row contractions (<= 64 splits)
    part_cap = std::max<int64_t>(256 * dims[0] * ldr, 64 * maxI * ldr);
    if (!mfma32) {
      if (need_tree && N >= 4) part_cap = std::max<int64_t>(part_cap, 64 * yl * ldr);
      T.alloc(h, C * ldr);
      if (need_tree) YL.alloc(h, yl * ldr);
    } else {
      part32_cap = 256 * dims[0] * rpf;
      if (need_tree && N >= 4) part32_cap = std::max<int64_t>(part32_cap, 64 * yl * rpf);
      part32.alloc(h, part32_cap);
      T32.alloc(h, std::max(C, dims[0]) * rpf);
      if (need_tree) YL32.alloc(h, yl * rpf);
    }
    part.alloc(h, part_cap);
  }
  // buffers of the paired left-node pass (dimension tree, FP32-output contractions)
  void enable_pair() {
    const int64_t yl = left_rows();
    YL2.alloc(h, yl * 2 * rpf);
    M2.alloc(h, 2 * dims[0] * ldr);          // [Mhat | M] of mode 0 from the paired pass
    if (N >= 4 && 64 * yl * 2 * rpf > part32_cap) {
      part32.~DevBuf(); new (&part32) DevBuf<float>();
      part32_cap = 64 * yl * 2 * rpf;
      part32.alloc(h, part32_cap);
    }
  }
  // the tree's root splits at half = ceil(N / 2) (DIM_HALF_SIZE, DimTrees.hpp:274): the left
  // node holds modes [0, half), the right node modes [half, N)
  int half() const { return (N + 1) / 2; }
  int64_t left_rows() const {
    int64_t r = 1;
    for (int k = 0; k < half(); ++k) r *= dims[k];
    return r;
  }
  double *F(int k) { return fac.p + off[k]; }
  double *Fb(double *base, int k) { return base + off[k]; }
  double *G(int k) { return grams.p + (int64_t)k * R * R; }
  const double *x64() const { return arith == PTN_ARITH_F64 ? h->x64 : nullptr; }
  const float *x32() const { return arith == PTN_ARITH_F64 ? nullptr : h->x32; }
  GemmMode gmode() const { return arith == PTN_ARITH_BF16X3 ? kGemmBf16x3 : kGemmFp32; }

  // Khatri-Rao operand of an FP32 contraction, packed in fragment order; with k2 the
  // operand carries both (2 rpf columns: k in the lower half, k2 in the upper).
  void pack(const KrSpec &k, int64_t K, int kperm, const KrSpec *k2 = nullptr, hipStream_t ps = nullptr) {
    const int width = k2 ? 2 * rpf : rpf;
    const int64_t frag = wpack_elems(K, width);
    if (frag > wcap) {
      HIPCHK(hipStreamSynchronize(h->stream));
      whi.~DevBuf(); new (&whi) DevBuf<uint16_t>();
      wlo.~DevBuf(); new (&wlo) DevBuf<uint16_t>();
      wf.~DevBuf(); new (&wf) DevBuf<float>();
      whi.alloc(h, frag * 8); wlo.alloc(h, frag * 8);
      wf.alloc(h, frag * 8);
      wcap = frag;
    }
    const bool x3 = arith == PTN_ARITH_BF16X3;
    pack_w(ps ? ps : h->stream, k, k2, K, width, ldr, kperm, x3 ? (void *)whi.p : nullptr,
           x3 ? (void *)wlo.p : nullptr, x3 ? nullptr : wf.p);
  }
};

// A contracted node: rows of R values, FP64 (ld = ldr) or FP32 (ld = rpf).  With sel
// (device int) the node starts sel_off values further when *sel == 1.
struct Node {
  const double *p64;
  const float *p32;
  int64_t ld;
  const int *sel = nullptr;
  int64_t sel_off = 0;
};

int pick_splits(int64_t blocks_per_split, int64_t work_per_split_unit, int64_t max_split) {
  int64_t s = (2048 + blocks_per_split - 1) / blocks_per_split;
  if (s > max_split) s = max_split;
  if (s > work_per_split_unit) s = work_per_split_unit;
  if (s < 1) s = 1;
  return (int)s;
}

// k splits of a strided contraction (gemm_cm) of Mrows rows over K.  FP32-stream launches run
// at most two blocks per CU (512 on the chip; bf16x3 by its padded LDS, FP32 products by their
// registers), so the split count is rounded down to give whole waves of 512 blocks: 200^4 at
// R = 64, 157 row blocks, 13 splits = 2041 blocks (bf16x3: 1.145 ms node with its reduction)
// against 14 = 2198 blocks (1.27 ms, a fifth wave 30% full).
int pick_splits_cm(const Work &w, int64_t Mrows, int64_t K);

KrSpec kr_of(Work &w, double *base, std::initializer_list<int> modes) {
  KrSpec k{};
  for (int m : modes) {
    k.dim[k.nm] = w.dims[m];
    k.fac[k.nm] = w.Fb(base, m);
    k.nm++;
  }
  return k;
}

// the Khatri-Rao operand of modes [lo, hi)
KrSpec kr_range(Work &w, double *base, int lo, int hi) {
  KrSpec k{};
  for (int m = lo; m < hi; ++m) {
    k.dim[k.nm] = w.dims[m];
    k.fac[k.nm] = w.Fb(base, m);
    k.nm++;
  }
  return k;
}

// out[m][r] = sum_k X[m + M k] W(k)[r]: into dst (S == 1) or via split partials.
// timed >= 0: the contraction kernel is bracketed by the events of pass `timed`
int pick_splits_cm(const Work &w, int64_t Mrows, int64_t K) {
  const int64_t blocks = (Mrows + 255) / 256;
  if (!w.mfma32) return pick_splits(blocks, K / 256, 64);
  int64_t s = 2048 / blocks;
  if (s > 64) s = 64;
  if (s > K / 256) s = K / 256;
  return s < 1 ? 1 : (int)s;
}

Node node_cm(Work &w, const KrSpec &k, int64_t Mrows, int64_t K, int S, double *dst64, float *dst32,
             int timed = -1) {
  ptn_handle h = w.h;
  hipStream_t s = h->stream;
  if (!w.mfma32) {
    w.pass_mark(timed, false, s);
    if (S == 1) gemm_cm_f64(s, w.x64(), w.x32(), Mrows, K, k, w.ldr, 1, dst64);
    else {
      REQUIRE(S * Mrows * w.ldr <= w.part_cap, PTN_ERR_INVALID, "partial buffer too small");
      gemm_cm_f64(s, w.x64(), w.x32(), Mrows, K, k, w.ldr, S, w.part);
    }
    w.pass_mark(timed, true, s);
    if (S > 1) reduce_splits(s, w.part, S, Mrows * w.ldr, dst64);
    return Node{dst64, nullptr, w.ldr};
  }
  w.pack(k, K, 0);
  w.pass_mark(timed, false, s);
  if (S == 1) gemm_cm_f32(s, w.gmode(), w.x32(), Mrows, K, w.rpf, 1, w.whi, w.wlo, w.wf, dst32);
  else {
    REQUIRE(S * Mrows * w.rpf <= w.part32_cap, PTN_ERR_INVALID, "partial buffer too small");
    gemm_cm_f32(s, w.gmode(), w.x32(), Mrows, K, w.rpf, S, w.whi, w.wlo, w.wf, w.part32);
  }
  w.pass_mark(timed, true, s);
  if (S > 1) reduce_splits_f32(s, w.part32, S, Mrows * w.rpf, dst32);
  return Node{nullptr, dst32, w.rpf};
}

// T[c][r] = sum_{a < L} X[a + L c] W(a)[r]
Node node_ck(Work &w, const KrSpec &k, int64_t L, int64_t C, hipStream_t s = nullptr, int timed = -1) {
  ptn_handle h = w.h;
  if (!s) s = h->stream;
  if (!w.mfma32) {
    w.pass_mark(timed, false, s);
    gemm_ck_f64(s, w.x64(), w.x32(), L, C, k, w.ldr, w.T);
    w.pass_mark(timed, true, s);
    return Node{w.T.p, nullptr, w.ldr};
  }
  w.pack(k, L, 1, nullptr, s);
  w.pass_mark(timed, false, s);
  gemm_ck_f32(s, w.gmode(), w.x32(), L, C, w.rpf, w.whi, w.wlo, w.wf, w.T32, 0, nullptr, w.ldr, nullptr);
  w.pass_mark(timed, true, s);
  return Node{nullptr, w.T32.p, w.rpf};
}

// Order-3 right node fused with its TTV: M[i2][r] = sum_{i0,i1} X(i0,i1,i2) A0[i0][r] A1[i1][r]
// in one pass over X (FP32-output kernels only).
void right_node_fused(Work &w, double *base, double *M) {
  ptn_handle h = w.h;
  const int64_t *d = w.dims;
  const int64_t wpf = (d[1] + 63) / 64;
  REQUIRE(d[2] * wpf * w.ldr <= w.part_cap, PTN_ERR_INVALID, "partial buffer too small");
  w.pack(kr_of(w, base, {0}), d[0], 1);
  w.pass_mark(1, false, h->stream);
  gemm_ck_f32(h->stream, w.gmode(), w.x32(), d[0], d[1] * d[2], w.rpf, w.whi, w.wlo, w.wf, nullptr, d[1],
              w.Fb(base, 1), w.ldr, w.part);
  w.pass_mark(1, true, h->stream);
  reduce_chunks(h->stream, w.part, d[2], wpf, w.ldr, M);
}

// M[i][r] = sum over the node's other digits of node[row][r] * (their factor rows)
void contract_node(Work &w, const Node &nd, const KrSpec &sp, int outmode, double *M) {
  ptn_handle h = w.h;
  hipStream_t s = h->stream;
  const int64_t In = sp.dim[outmode];
  int64_t Cother = 1;
  for (int j = 0; j < sp.nm; ++j)
    if (j != outmode) Cother *= sp.dim[j];
  if (sp.nm == 2) {
    // two digits: coalesced row-block kernel (one output digit, one contracted digit)
    const int64_t D0 = sp.dim[0], D1 = sp.dim[1];
    int S = outmode == 0 ? pick_splits((D0 + 7) / 8, D1 / 16, 64) : pick_splits(D1, D0 / 64, 64);
    // one block per output row already fills the device: no split, M written directly
    if (outmode == 1 && D1 >= 512) S = 1;
    REQUIRE(S * In * w.ldr <= w.part_cap, PTN_ERR_INVALID, "partial buffer too small");
    contract2(s, nd.p64, nd.p32, nd.ld, D0, D1, outmode, sp.fac[1 - outmode], w.ldr, w.R, S, S == 1 ? M : w.part,
              nd.sel, nd.sel_off);
    if (S > 1) reduce_splits(s, w.part, S, In * w.ldr, M);
    return;
  }
  const int S = pick_splits(In, Cother / 64, 64);
  REQUIRE(S * In * w.ldr <= w.part_cap, PTN_ERR_INVALID, "partial buffer too small");
  contract_rows(s, nd.p64, nd.p32, nd.ld, sp, outmode, w.ldr, S, w.part, nd.sel, nd.sel_off);
  reduce_splits(s, w.part, S, In * w.ldr, M);
}


// M (row-major I_mode x ldr) = X_(mode) * KR(base factors), standard form.
void mttkrp_standard(Work &w, double *base, int mode, double *M) {
  const int N = w.N;
  const int64_t I0 = w.dims[0];
  if (mode == 0) {
    KrSpec k{};
    for (int j = 1; j < N; ++j) { k.dim[k.nm] = w.dims[j]; k.fac[k.nm] = w.Fb(base, j); k.nm++; }
    const int64_t Ktot = w.total / I0;
    const int S = pick_splits((I0 + 63) / 64, Ktot / 256, 256);
    if (!w.mfma32) {
      gemm_cm_f64(w.h->stream, w.x64(), w.x32(), I0, Ktot, k, w.ldr, S, w.part);
      reduce_splits(w.h->stream, w.part, S, I0 * w.ldr, M);
    } else {
      // FP32 partials summed in FP64 by the row contraction of a 1-digit node per split
      Node nd = node_cm(w, k, I0, Ktot, S, nullptr, w.T32);
      KrSpec sp{};
      sp.nm = 1; sp.dim[0] = I0;
      contract_node(w, nd, sp, 0, M);
    }
    return;
  }
  if (w.mfma32 && N == 3 && mode == 2) {
    right_node_fused(w, base, M);
    return;
  }
  Node nd = node_ck(w, kr_of(w, base, {0}), I0, w.total / I0);
  KrSpec sp{};
  for (int j = 1; j < N; ++j) { sp.dim[sp.nm] = w.dims[j]; sp.fac[sp.nm] = w.Fb(base, j); sp.nm++; }
  contract_node(w, nd, sp, mode - 1, M);
}

// Dimension-tree MTTKRP of `mode` (tree of ExprTree::Create, DimTrees.hpp:541-628),
// recomputing the node it needs where UpdateTree (DimTrees.hpp:862-928) would in an
// ALS sweep: the left node at mode 0, the right node at its first leaf.  N = 4 nodes
// are fused GEMM + TTV contractions (one pass over X each).
void mttkrp_tree(Work &w, double *base, int mode, double *M) {
  const int N = w.N;
  const int64_t *d = w.dims;
  auto keep = [](Work::NodeRef &dst, const Node &n) {
    dst.p64 = n.p64; dst.p32 = n.p32; dst.ld = n.ld; dst.sel = nullptr; dst.sel_off = 0;
  };
  auto node = [](const Work::NodeRef &r) { return Node{r.p64, r.p32, r.ld, r.sel, r.sel_off}; };
  if (N == 3) {
    if (mode == 0 && !w.left_ready)
      keep(w.left, node_cm(w, kr_of(w, base, {2}), d[0] * d[1], d[2], 1, w.YL.p, w.YL32.p, 0));
    w.left_ready = false;
    if (mode <= 1) {
      if (mode == 0 && w.m0_ready) {         // chosen with the line-search decision
        w.m0_ready = false;
        return;
      }
      contract_node(w, node(w.left), kr_of(w, base, {0, 1}), mode, M);
      return;
    }
    if (w.right_ready) {                     // formed on the aux stream after mode 0 (right_node_early)
      w.right_ready = false;
      ptn_handle h = w.h;
      HIPCHK(hipStreamWaitEvent(h->stream, h->ev[3], 0));
      contract_node(w, node(w.right), kr_of(w, base, {1, 2}), 1, M);
      return;
    }
    if (w.mfma32) {
      right_node_fused(w, base, M);
      return;
    }
    keep(w.right, node_ck(w, kr_of(w, base, {0}), d[0], d[1] * d[2], nullptr, 1));
    contract_node(w, node(w.right), kr_of(w, base, {1, 2}), 1, M);
    return;
  }
  // order >= 4: the root's two children, each one pass over X (the GEMM node with the
  // Khatri-Rao rows formed on the fly, so the root-to-child TTV chain is fused); every leaf
  // contracts its child over the child's other modes
  const int hf = w.half();
  const int64_t Ml = w.left_rows(), Kl = w.total / Ml;
  if (mode == 0 && !w.left_ready) {
    const int S = pick_splits_cm(w, Ml, Kl);
    keep(w.left, node_cm(w, kr_range(w, base, hf, N), Ml, Kl, S, w.YL.p, w.YL32.p, 0));
  }
  w.left_ready = false;
  if (mode == 0 && w.m0_ready) {             // chosen with the line-search decision
    w.m0_ready = false;
    return;
  }
  if (mode == hf) keep(w.right, node_ck(w, kr_range(w, base, 0, hf), Ml, Kl, nullptr, 1));
  if (mode < hf) contract_node(w, node(w.left), kr_range(w, base, 0, hf), mode, M);
  else contract_node(w, node(w.right), kr_range(w, base, hf, N), mode - hf, M);
}

// The order-3 tree's right node X x_0 A_0 (R x I1 x I2, FP32) depends only on A_0, so in
// a sweep it is formed on the aux stream as soon as mode 0 is updated, beside mode 1's
// contraction, eigen set-up and Nesterov solve; mode 2 then contracts it with A_1
// (DimTrees.hpp:668-800: the same node, formed earlier).  The main stream waits for it
// at mode 2.  `after`: an event the caller recorded on the main stream with nothing
// enqueued there since (the aux stream then waits on it and the main queue takes no
// second record); null records one here.
void right_node_early(Work &w, double *base, hipEvent_t after = nullptr) {
  ptn_handle h = w.h;
  if (!after) {
    HIPCHK(hipEventRecord(h->ev[2], h->stream));
    after = h->ev[2];
  }
  HIPCHK(hipStreamWaitEvent(h->aux, after, 0));
  const Node nd = node_ck(w, kr_of(w, base, {0}), w.dims[0], w.dims[1] * w.dims[2], h->aux, 1);
  w.right.p64 = nd.p64; w.right.p32 = nd.p32; w.right.ld = nd.ld; w.right.sel = nullptr; w.right.sel_off = 0;
  HIPCHK(hipEventRecord(h->ev[3], h->aux));
  w.right_ready = true;
}

// Line-search pass fused with the next sweep's left node: one pass over X forms
// YL2[m] = [X x_rest A_rest | X x_rest Ahat_rest] (rest = modes 2.. of the current and of
// the accelerated factors, 2 rpf columns per row), then one launch contracts both halves
// over i1: M2 = [Mhat | M], Mhat (I_0 x ldr) = sum_{i1} YL2hat[i0 + I0 i1] Ahat_1[i1], so
// <Mhat, Ahat_0> is the fit term <X, [[Ahat]]> of accel_cost_function
// (CpdDimTree.hpp:174-181) summed in another order, and M the mode-0 MTTKRP of the current
// factors (the leaf of the left node, as mttkrp_tree forms it).  The decision copies one of
// the two into w.M.
void left_node_pair(Work &w, double *cur, double *acc) {
  ptn_handle h = w.h;
  hipStream_t s = h->stream;
  const bool n3 = w.N == 3;
  const int hf = w.half();
  const KrSpec k1 = kr_range(w, cur, hf, w.N);
  const KrSpec k2 = kr_range(w, acc, hf, w.N);
  const int64_t Mrows = w.left_rows(), K = w.total / Mrows;
  const int width = 2 * w.rpf;
  w.pack(k1, K, 0, &k2);
  const int S = n3 ? 1 : pick_splits_cm(w, Mrows, K);
  w.pass_mark(0, false, s);
  if (S == 1) {
    gemm_cm_f32(s, w.gmode(), w.x32(), Mrows, K, width, 1, w.whi, w.wlo, w.wf, w.YL2);
    w.pass_mark(0, true, s);
  } else {
    REQUIRE(S * Mrows * width <= w.part32_cap, PTN_ERR_INVALID, "partial buffer too small");
    gemm_cm_f32(s, w.gmode(), w.x32(), Mrows, K, width, S, w.whi, w.wlo, w.wf, w.part32);
    w.pass_mark(0, true, s);
    reduce_splits_f32(s, w.part32, S, Mrows * width, w.YL2);
  }
  const KrSpec la = kr_range(w, acc, 0, hf), lc = kr_range(w, cur, 0, hf);
  const int64_t D0 = la.dim[0], D1 = la.dim[1];
  if (la.nm == 2) {
    const int S2 = pick_splits((D0 + 7) / 8, D1 / 16, 64);   // contract_node's split of outmode 0
    REQUIRE(2 * S2 * D0 * w.ldr <= w.part_cap, PTN_ERR_INVALID, "partial buffer too small");
    contract2_pair(s, w.YL2.p + w.rpf, w.YL2.p, width, D0, D1, 0, la.fac[1], lc.fac[1], w.ldr, w.R, S2, w.part);
    reduce_splits(s, w.part, S2, 2 * D0 * w.ldr, w.M2);
  } else {                                   // three or more digits (orders 5 .. 8): one contraction per half
    contract_node(w, Node{nullptr, w.YL2.p + w.rpf, width}, la, 0, w.M2);
    contract_node(w, Node{nullptr, w.YL2.p, width}, lc, 0, w.M2 + D0 * w.ldr);
  }
  w.m0_ready = true;
}

// Hadamard order of V_n: ascending (PartialCwiseProd, right-nested) for the standard
// form; the tree path order (TreeMode_N_Product / TTVs_util) for the dimension tree.
int v_order(int N, int mode, bool tree, int *list) {
  if (tree && N == 3) {
    static const int o3[3][2] = {{2, 1}, {2, 0}, {0, 1}};
    list[0] = o3[mode][0]; list[1] = o3[mode][1];
    return 2;
  }
  if (tree && N == 4) {
    static const int o4[4][3] = {{3, 2, 1}, {3, 2, 0}, {0, 1, 3}, {0, 1, 2}};
    for (int j = 0; j < 3; ++j) list[j] = o4[mode][j];
    return 3;
  }
  int n = 0;
  for (int k = 0; k < N; ++k)
    if (k != mode) list[n++] = k;
  return n;
}

// Inputs of one NNLS update: V_n (PartialCwiseProd of the Grams), its extreme
// eigenvalues with the GLambda shift, and the momentum coefficients.  Slot n serves
// mode n; slot N holds the mode-0 set-up for the accelerated factors while the line
// search is undecided.
struct ModeIn {
  double *V;
  NnlsParams *p;
  double *betas;
};

ModeIn slot(Work &w, int s) {
  return ModeIn{w.Vm.p + (int64_t)s * w.R * w.R, w.prm.p + s, w.betas.p + (int64_t)s * (kMaxInner + 1)};
}

void prep_mode(Work &w, const double *gbase, int mode, bool tree, const ModeIn &in, int max_inner,
               hipStream_t s) {
  int list[kMaxOrder];
  int n = v_order(w.N, mode, tree, list);
  const double *gp[kMaxOrder];
  for (int j = 0; j < n; ++j) gp[j] = gbase + (int64_t)list[j] * w.R * w.R;
  nnls_setup(s, w.R, n, gp, in.V, in.p, max_inner, in.betas);
}

void gram_of(Work &w, const double *A, int64_t I, double *G) { gram(w.h->stream, I, w.R, w.ldr, A, G, w.gpart); }

double read_scalar(ptn_handle h, const double *dev) {
  double v;
  HIPCHK(hipMemcpyAsync(&v, dev, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return v;
}

void frob_norm(ptn_handle h) {
  if (h->frob_valid) return;
  square_norm_partials(h->stream, h->total, h->x64, h->x64 ? nullptr : h->x32, h->part, kNormParts);
  sum_partials(h->stream, h->part, kNormParts, h->scal);
  h->frob_host = read_scalar(h, h->scal);
  h->frob_valid = true;
}

void upload_factors(Work &w, const double *packed_host, double *dst_rows) {
  ptn_handle h = w.h;
  int64_t n = 0;
  for (int k = 0; k < w.N; ++k) n += w.dims[k] * w.R;
  DevBuf<double> tmp(h, n);
  HIPCHK(hipMemcpyAsync(tmp.p, packed_host, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  int64_t o = 0;
  for (int k = 0; k < w.N; ++k) {
    colmajor_to_rows(h->stream, w.dims[k], w.R, w.ldr, tmp.p + o, dst_rows + w.off[k]);
    o += w.dims[k] * w.R;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
}

void download_factors(Work &w, const double *src_rows, double *packed_host) {
  ptn_handle h = w.h;
  int64_t n = 0;
  for (int k = 0; k < w.N; ++k) n += w.dims[k] * w.R;
  DevBuf<double> tmp(h, n);
  int64_t o = 0;
  for (int k = 0; k < w.N; ++k) {
    rows_to_colmajor(h->stream, w.dims[k], w.R, w.ldr, src_rows + w.off[k], tmp.p + o);
    o += w.dims[k] * w.R;
  }
  HIPCHK(hipMemcpyAsync(packed_host, tmp.p, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
}

double cost_explicit(Work &w, const double *base) {
  ptn_handle h = w.h;
  const double *fp[kMaxOrder];
  for (int k = 0; k < w.N; ++k) fp[k] = base + w.off[k];
  cost_explicit_partials(h->stream, w.N, w.dims, w.R, w.ldr, fp, h->x64, h->x64 ? nullptr : h->x32, w.part,
                         kNormParts);
  sum_partials(h->stream, w.part, kNormParts, w.scal.p + 3);
  return std::sqrt(read_scalar(h, w.scal.p + 3));
}

// update_factor, Cpd.hpp:189-228, nonnegativity branch on the device (V_n, its
// eigenvalues and the momentum coefficients prepared in `in`)
void update_nonneg(Work &w, int mode, const ptn_options &o, const ModeIn &in) {
  ptn_handle h = w.h;
  hipStream_t s = h->stream;
  const int64_t I = w.dims[mode];
  double *A = w.F(mode);
  NnlsParams *p = in.p;
  const int64_t need = (int64_t)(o.max_inner + 1) * w.maxI * w.ldr;
  if (need > w.hist_cap) {
    HIPCHK(hipStreamSynchronize(s));
    w.hist.~DevBuf();
    new (&w.hist) DevBuf<double>();
    w.hist.alloc(h, need);
    w.hist_cap = need;
  }
  const bool warm = o.warm_start != 0, harvest = o.mlp_harvest != 0;
  MlpState *ms = (warm || harvest) ? &mlp_of(h) : nullptr;
  if (warm || harvest) HIPCHK(hipMemcpyAsync(w.Aold, A, I * w.ldr * sizeof(double), hipMemcpyDeviceToDevice, s));
  const double *start = nullptr;
  if (warm) {
    // the network proposes each row's start; rows failing the KKT test keep A (k_warm_select)
    mlp_features(s, I, w.R, w.ldr, in.V, w.M, A, nullptr, ms->X.p, nullptr, nullptr, ms->scale.p);
    ms->forward(s, (int)I);
    if (o.warm_refine > 0) {
      // every finite proposal first takes warm_refine Nesterov steps of the same subproblem
      // (anchor A, the global stop test included); the test then judges the refined rows
      warm_select(s, I, w.R, w.ldr, in.V, p, w.M, A, ms->Y.p, ms->scale.p, nullptr, INFINITY, ms->start.p,
                  nullptr);
      nnls_iterate(s, I, w.R, w.ldr, in.V, p, in.betas, w.M, A, ms->start.p, o.nesterov_delta_1,
                   o.nesterov_delta_2, o.warm_refine, w.mask, w.hist, w.colsum);
      nnls_select(s, I, w.R, w.ldr, w.hist, w.mask, o.warm_refine, ms->start.p, w.colsum, p);
      warm_select(s, I, w.R, w.ldr, in.V, p, w.M, A, nullptr, nullptr, ms->start.p, o.warm_kkt_tol, ms->start.p,
                  &p->warm_rows);
    } else {
      warm_select(s, I, w.R, w.ldr, in.V, p, w.M, A, ms->Y.p, ms->scale.p, nullptr, o.warm_kkt_tol, ms->start.p,
                  &p->warm_rows);
    }
    start = ms->start.p;
  }
  // without a warm start the start point A_old stays in hist[0] for the fallback
  nnls_iterate(s, I, w.R, w.ldr, in.V, p, in.betas, w.M, A, start, o.nesterov_delta_1, o.nesterov_delta_2,
               o.max_inner, w.mask, w.hist, w.colsum);
  nnls_finish(s, I, w.R, w.ldr, w.hist, w.mask, o.max_inner, A, w.colsum, warm ? w.Aold.p : w.hist.p, w.G(mode),
              w.gpart, p);
  if (harvest) {
    // (V, M row, previous row) -> the solution row, appended to the training set
    mlp_features(s, I, w.R, w.ldr, in.V, w.M, w.Aold, A, ms->dx.p + ms->n * kMlpIn, ms->dt.p + ms->n * kMlpR,
                 ms->drank.p + ms->n, nullptr);
    ms->n += I;
    ms->hrank.insert(ms->hrank.end(), (size_t)I, (unsigned char)w.R);
  }
}

// update_factor, Cpd.hpp:189-228: the constraint of the mode picks the update; G_n =
// A_n^T A_n follows (Cpd.hpp:227)
void update_factor(Work &w, int mode, const ptn_options &o, const ModeIn &in) {
  ptn_handle h = w.h;
  hipStream_t s = h->stream;
  const int64_t I = w.dims[mode];
  const int R = w.R;
  double *A = w.F(mode);
  double *c0 = w.cscr.p, *c1 = c0 + (int64_t)R * R, *c2 = c1 + (int64_t)R * R;
  switch (o.constraints[mode]) {
    case PTN_NONNEGATIVITY:
      update_nonneg(w, mode, o, in);
      return;
    case PTN_UNCONSTRAINED:                   // A = M V^{-1} (Cpd.hpp:196-200)
      lu_inverse(s, R, in.V, c0);
      apply_right(s, I, R, w.ldr, w.M, c0, A);
      break;
    case PTN_ORTHOGONALITY:                   // A = M (M^T M)^{-1/2} (Cpd.hpp:210-219)
      gram_of(w, w.M, I, c0);
      jacobi_eig(s, R, c0, c2, c1);
      inv_sqrt_matrix(s, R, c1, c2, c0);
      apply_right(s, I, R, w.ldr, w.M, c0, A);
      break;
    default:                                  // sparsity / constant: the factor is left as it is
      break;
  }
  params_no_nnls(s, in.p);
  gram_of(w, A, I, w.G(mode));
}

}  // namespace

// ======================================================================================
extern "C" {

int ptn_abi_version(void) { return 1; }

void ptn_default_options(ptn_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_iter = 20;
  o->threshold_error = 1e-3;
  o->nesterov_delta_1 = 1e-2;
  o->nesterov_delta_2 = 1e-2;
  o->accel_coeff = 3;
  o->accel_fail = 0;
  o->acceleration = 1;
  o->normalization = 1;
  for (int i = 0; i < PTN_MAX_ORDER; ++i) o->constraints[i] = PTN_NONNEGATIVITY;
  o->mttkrp_form = PTN_MTTKRP_STANDARD;
  o->arith = PTN_ARITH_F64;
  o->max_inner = 50;
  o->fuse_accel = 0;
  o->warm_start = 0;
  o->warm_kkt_tol = 1e-6;
  o->warm_refine = 0;
  o->mlp_harvest = 0;
}

ptn_status ptn_create(ptn_handle *out, int device) {
  if (!out) return PTN_ERR_INVALID;
  *out = nullptr;
  ptn_handle h = new ptn_handle_s();
  h->device = device;
  ptn_status st = run(h, [&] {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    REQUIRE(device >= 0 && device < n, PTN_ERR_INVALID, "no such device");
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
    for (hipEvent_t &e : h->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipMalloc((void **)&h->scal, 16 * sizeof(double)));
    HIPCHK(hipMalloc((void **)&h->part, kNormParts * sizeof(double)));
    REQUIRE(device < kNoallocMaxDevices, PTN_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[device]) {
      void *ptr = nullptr;
      HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      REQUIRE(ptr != nullptr, PTN_ERR_HIP, "no device address for the no-allocation scratch");
      noalloc_ptr[device] = ptr;
    }
  });
  if (st != PTN_OK) {
    ptn_destroy(h);
    return st;
  }
  *out = h;
  return PTN_OK;
}

void ptn_destroy(ptn_handle h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->side) (void)hipStreamSynchronize(h->side);
  if (h->aux) (void)hipStreamSynchronize(h->aux);
  if (h->x64) (void)hipFree(h->x64);
  if (h->x32) (void)hipFree(h->x32);
  delete h->batch;
  delete h->mlp;
  if (h->scal) (void)hipFree(h->scal);
  if (h->part) (void)hipFree(h->part);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->aux) (void)hipStreamDestroy(h->aux);
  for (hipEvent_t e : h->ev)
    if (e) (void)hipEventDestroy(e);
  delete h;
}

const char *ptn_last_error(ptn_handle h) { return h ? h->err.c_str() : "null handle"; }

}  // extern "C"

namespace {

ptn_status run(ptn_handle h, const std::function<void()> &fn) {
  if (!h) return PTN_ERR_INVALID;
  try {
    if (h->stream) {
      hipError_t e = hipSetDevice(h->device);
      if (e != hipSuccess) {
        h->err = hipGetErrorString(e);
        return PTN_ERR_HIP;
      }
    }
    fn();
    return PTN_OK;
  } catch (const Fail &f) {
    return f.code;
  } catch (const std::bad_alloc &) {
    h->err = "host allocation failed";
    return PTN_ERR_OOM;
  } catch (...) {
    h->err = "unexpected exception";
    return PTN_ERR_HIP;
  }
}

void alloc_tensor(ptn_handle h, int N, const int64_t *dims, int storage) {
  REQUIRE(storage >= 1 && storage <= 3, PTN_ERR_INVALID, "storage must be F64, F32 or BOTH");
  check_dims(h, N, dims);
  if (h->x64) { HIPCHK(hipFree(h->x64)); h->x64 = nullptr; }
  if (h->x32) { HIPCHK(hipFree(h->x32)); h->x32 = nullptr; }
  h->N = N;
  for (int k = 0; k < N; ++k) h->dims[k] = dims[k];
  h->total = prod(dims, N);
  h->frob_valid = false;
  // kTensorPad zero elements after the tensor: the contiguous-K contractions read whole
  // 32-element steps, which may run past the end of the last fiber
  if (storage & PTN_TENSOR_F64) {
    HIPCHK(hipMalloc((void **)&h->x64, (h->total + kTensorPad) * sizeof(double)));
    HIPCHK(hipMemsetAsync(h->x64 + h->total, 0, kTensorPad * sizeof(double), h->stream));
  }
  if (storage & PTN_TENSOR_F32) {
    HIPCHK(hipMalloc((void **)&h->x32, (h->total + kTensorPad) * sizeof(float)));
    HIPCHK(hipMemsetAsync(h->x32 + h->total, 0, kTensorPad * sizeof(float), h->stream));
  }
}

void check_arith(ptn_handle h, int arith) {
  REQUIRE(arith >= PTN_ARITH_F64 && arith <= PTN_ARITH_F32_F64, PTN_ERR_INVALID, "bad arith");
  REQUIRE(h->N > 0, PTN_ERR_STATE, "no tensor set on the handle");
  if (arith == PTN_ARITH_F64) REQUIRE(h->x64, PTN_ERR_STATE, "FP64 arithmetic needs the FP64 tensor copy");
  else REQUIRE(h->x32, PTN_ERR_STATE, "FP32 tensor streams need the FP32 tensor copy");
}

}  // namespace

extern "C" {

ptn_status ptn_set_tensor(ptn_handle h, int N, const int64_t *dims, const double *X, int storage) {
  return run(h, [&] {
    REQUIRE(X != nullptr, PTN_ERR_INVALID, "X is null");
    alloc_tensor(h, N, dims, storage);
    if (h->x64) {
      HIPCHK(hipMemcpyAsync(h->x64, X, h->total * sizeof(double), hipMemcpyHostToDevice, h->stream));
      if (h->x32) cast_f64_to_f32(h->stream, h->total, h->x64, h->x32);
    } else {
      DevBuf<double> tmp(h, h->total);
      HIPCHK(hipMemcpyAsync(tmp.p, X, h->total * sizeof(double), hipMemcpyHostToDevice, h->stream));
      cast_f64_to_f32(h->stream, h->total, tmp.p, h->x32);
      HIPCHK(hipStreamSynchronize(h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

ptn_status ptn_generate_tensor(ptn_handle h, int N, const int64_t *dims, int R_true, uint64_t seed, int dist,
                               double snr_db, int storage) {
  return run(h, [&] {
    REQUIRE(R_true >= 1 && R_true <= 4096, PTN_ERR_INVALID, "bad true rank");
    REQUIRE(dist == 0 || dist == 1, PTN_ERR_INVALID, "dist must be 0 (uniform) or 1 (exponential)");
    alloc_tensor(h, N, dims, storage);
    int64_t nf = 0;
    for (int k = 0; k < N; ++k) nf += dims[k] * R_true;
    DevBuf<double> fac(h, nf);
    gen_factors(h->stream, N, dims, R_true, seed, dist, fac.p);
    cp_full(h->stream, N, dims, R_true, fac.p, h->x64, h->x32);
    if (!std::isnan(snr_db)) {
      square_norm_partials(h->stream, h->total, h->x64, h->x64 ? nullptr : h->x32, h->part, kNormParts);
      sum_partials(h->stream, h->part, kNormParts, h->scal + 1);
      gen_noise_norm(h->stream, h->total, seed, h->part, kNormParts);
      sum_partials(h->stream, h->part, kNormParts, h->scal + 2);
      noise_scale(h->stream, h->scal + 1, h->scal + 2, snr_db, h->scal + 3);
      add_noise(h->stream, h->total, seed, h->scal + 3, h->x64, h->x32);
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
  });
}

ptn_status ptn_generate_factors(ptn_handle h, int N, const int64_t *dims, int R, uint64_t seed, int dist,
                                double *packed_out) {
  return run(h, [&] {
    check_dims(h, N, dims);
    REQUIRE(packed_out, PTN_ERR_INVALID, "output is null");
    int64_t nf = 0;
    for (int k = 0; k < N; ++k) nf += dims[k] * R;
    DevBuf<double> fac(h, nf);
    gen_factors(h->stream, N, dims, R, seed, dist, fac.p);
    HIPCHK(hipMemcpyAsync(packed_out, fac.p, nf * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

ptn_status ptn_tensor_norm_sq(ptn_handle h, double *out) {
  return run(h, [&] {
    REQUIRE(h->N > 0, PTN_ERR_STATE, "no tensor set on the handle");
    frob_norm(h);
    *out = h->frob_host;
  });
}

ptn_status ptn_cpd(ptn_handle h, int R, const double *init_packed, const ptn_options *opt, double *out_packed,
                   ptn_result *res, ptn_sweep_log *log, int log_cap) {
  return run(h, [&] {
    REQUIRE(opt && init_packed && out_packed && res, PTN_ERR_INVALID, "null argument");
    const ptn_options &o = *opt;
    check_arith(h, o.arith);
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(o.max_iter >= 1, PTN_ERR_INVALID, "max_iter must be >= 1");
    REQUIRE(o.max_inner >= 0 && o.max_inner <= kMaxInner, PTN_ERR_UNSUPPORTED, "max_inner must be in [0, 255]");
    REQUIRE(o.accel_coeff >= 1, PTN_ERR_INVALID, "accel_coeff must be >= 1");
    const bool tree = o.mttkrp_form == PTN_MTTKRP_DIMTREE;
    REQUIRE(o.mttkrp_form == PTN_MTTKRP_STANDARD || tree, PTN_ERR_INVALID, "bad mttkrp_form");
    REQUIRE(!tree || h->N <= kMaxOrder, PTN_ERR_UNSUPPORTED, "dimension tree built for N = 3 .. 8");
    for (int k = 0; k < h->N; ++k)
      REQUIRE(o.constraints[k] >= PTN_UNCONSTRAINED && o.constraints[k] <= PTN_CONSTANT, PTN_ERR_INVALID,
              "bad constraint");
    const bool pair = o.fuse_accel != 0;
    REQUIRE(!pair || (tree && (o.arith == PTN_ARITH_F32 || o.arith == PTN_ARITH_BF16X3)), PTN_ERR_UNSUPPORTED,
            "fuse_accel needs the dimension tree with F32 or BF16X3 contractions");
    REQUIRE(!o.warm_start || (h->mlp && h->mlp->ready), PTN_ERR_STATE,
            "warm_start needs a network (ptn_mlp_init or ptn_mlp_set_params)");
    REQUIRE(o.warm_refine >= 0 && o.warm_refine <= o.max_inner, PTN_ERR_INVALID,
            "warm_refine must be in 0 .. max_inner");

    Work w(h, R, o.arith, tree);
    if (o.warm_start || o.mlp_harvest) {
      MlpState &mst = mlp_of(h);
      mst.ensure_rows((int)w.maxI);
      if (o.mlp_harvest) {
        int64_t per = 0;
        for (int k = 0; k < w.N; ++k) per += w.dims[k];
        mst.ensure_data(mst.n + per * o.max_iter);
      }
    }
    if (pair && o.acceleration) w.enable_pair();
    const int N = w.N, last = N - 1;
    upload_factors(w, init_packed, w.fac);
    for (int k = 0; k < N; ++k) gram_of(w, w.F(k), w.dims[k], w.G(k));

    // choose_normilization_factor, Normalize.hpp:53-67
    int wf = 0;
    bool all_orth = true;
    if (o.normalization)
      for (int k = 0; k < N; ++k)
        if (o.constraints[k] != PTN_ORTHOGONALITY) { all_orth = false; wf = k; break; }

    frob_norm(h);
    ptn_result r{};
    r.frob_sq = h->frob_host;
    r.f_value = cost_explicit(w, w.fac);   // cost_function_init
    r.rel_cost = r.f_value / std::sqrt(r.frob_sq);
    r.accel_coeff = o.accel_coeff;
    r.accel_fail = o.accel_fail;
    HIPCHK(hipMemcpyAsync(w.scal.p, &r.frob_sq, sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));

    // The side stream forms V_n, its eigenvalues and the momentum coefficients of mode n
    // while the main stream contracts the tensor for mode n; during the line-search pass
    // it also forms the accelerated factors' Grams, V_hat, the sweep's cost and both
    // mode-0 set-ups of the next sweep.  The aux stream forms the order-3 tree's right node
    // (right_node_early).  All three run at the default priority.
    hipStream_t ms = h->stream, ss = h->side;
    auto fork = [&] {      // side waits for the main stream's work so far
      HIPCHK(hipEventRecord(h->ev[0], ms));
      HIPCHK(hipStreamWaitEvent(ss, h->ev[0], 0));
    };
    auto join = [&] {      // main waits for the side stream's work so far
      HIPCHK(hipEventRecord(h->ev[1], ss));
      HIPCHK(hipStreamWaitEvent(ms, h->ev[1], 0));
    };
    ModeIn sel[kMaxOrder];
    bool prepped0 = false;
    // The host enqueues sweep it + 1 before it reads sweep it back (page-locked buffers
    // the kernels write directly; the last of them publishes the sweep index), so the
    // device never waits for the host.  The line-search
    // decision and its counters live on the device (flags[0..3]); a sweep that ends the
    // run on the threshold is found one sweep late, and the factors come from its
    // snapshot.
    struct Readback {
      NnlsParams hp[kMaxOrder + 1];
      double f, fa;
      int flags[4];
      unsigned seq;                           // the sweep index, published last by the device
    };
    PinnedBuf<Readback> rbuf0(h), rbuf1(h);
    struct PassEvents {                       // time_passes: 2 sweep slots x (start, end) x 2 passes
      hipEvent_t e[2][4] = {};
      ~PassEvents() {
        for (auto &r : e)
          for (hipEvent_t x : r)
            if (x) (void)hipEventDestroy(x);
      }
    } pev;
    if (o.time_passes) {
      for (auto &r : pev.e)
        for (hipEvent_t &x : r) HIPCHK(hipEventCreate(&x));
      w.kev = pev.e;
    }
    Readback *rbs[2] = {rbuf0.p, rbuf1.p};
    Readback *rbd[2] = {rbuf0.dev, rbuf1.dev};     // the same buffers as the device addresses them
    DevBuf<double> snap0(h, w.fsz), snap1(h, w.fsz);
    double *snaps[2] = {snap0.p, snap1.p};
    EventPair evs(h);
    {
      const int fl[4] = {0, 0, o.accel_fail, o.accel_coeff};
      HIPCHK(hipMemcpy(w.flags.p, fl, sizeof(fl), hipMemcpyHostToDevice));
    }
    int c_known = o.accel_coeff;     // coefficient after the last sweep read back
    std::chrono::steady_clock::time_point t0;
    bool tried_at[2] = {false, false};

    auto enqueue = [&](int it) {
      Readback *rbdev = rbd[it & 1];
      w.tslot = w.kev ? (it & 1) : -1;
      // termination (Cpd.hpp:339-344): the iteration cap is known here; the threshold
      // test is repeated on the device, which then takes no step
      const bool cap = it >= (int)o.max_iter;
      const bool step_now = !cap && o.acceleration && it > 1;
      // with the paired pass the sweep's cost is needed first by the line-search
      // decision, so it runs on the side stream beside the next pass (Normalize leaves
      // the last factor and its Gram as they are unless it is the weight factor)
      const bool side_cost = step_now && pair && !(o.normalization && !all_orth && wf == last);
      const bool early_right = tree && w.mfma32 && N == 3;
      for (int n = 0; n < N; ++n) {
        if (!(n == 0 && prepped0)) {
          sel[n] = slot(w, n);
          fork();
          prep_mode(w, w.grams, n, tree, sel[n], o.max_inner, ss);
        }
        // mode 1's set-up is queued on the side stream before the right node on the third,
        // so that its eigen wave starts before the right node's blocks fill the CUs; the
        // third stream waits on the fork's event (only the side stream was fed since)
        if (n == 1 && early_right) right_node_early(w, w.fac, h->ev[0]);
        // mode 2's set-up ends on the side stream before the right node on the third, so
        // the main stream joins the side stream before it waits for the right node: both
        // waits then pass before the leaf contraction instead of one on each side of it
        const bool join_early = n == 2 && early_right;
        if (join_early) join();
        if (tree) mttkrp_tree(w, w.fac, n, w.M);
        else mttkrp_standard(w, w.fac, n, w.M);
        if (!(n == 0 && prepped0) && !join_early) join();   // mode 0: joined after the line-search pass
        update_factor(w, n, o, sel[n]);
        if (n == last && !side_cost)
          cost_gram(ms, w.dims[last] * w.ldr, w.M, w.F(last), R, sel[last].V, w.G(last), w.scal.p, w.scal.p + 1, 0);
      }
      prepped0 = false;
      ExportList ex{};                 // the sweep's readback: NNLS parameters and the cost
      ex.src[0] = (const unsigned *)w.prm.p; ex.dst[0] = (unsigned *)rbdev->hp;
      ex.words[0] = (int)(sizeof(NnlsParams) * (N + 1) / 4);
      ex.src[1] = (const unsigned *)(w.scal.p + 1); ex.dst[1] = (unsigned *)&rbdev->f; ex.words[1] = 2;
      ex.count = side_cost ? 1 : 2;    // with side_cost the decision exports the cost
      // the last kernel of the sweep that writes the readback publishes the sweep index
      // (untimed runs: the host waits for it instead of an event at the end of the sweep)
      const bool will_try = !cap && o.acceleration && it > 1;
      if (!will_try && !w.kev) { ex.seq_dst = &rbdev->seq; ex.seq = (unsigned)it; }
      // line_search_accel, Cpd.hpp:247-278: step (k+1)^(1/accel_coeff); the device's
      // coefficient is c_known or, after one more fifth failure, c_known + 1.  A_hat =
      // A_old + step (A - A_old); A_old = A (the post-Normalize factors) and the sweep's
      // snapshot in the same pass (Normalize's row pass when it runs)
      AxpyArgs ax{};
      if (step_now) {
        ax.step_c = std::pow((double)(it + 1), 1.0 / (double)c_known);
        ax.step_c1 = std::pow((double)(it + 1), 1.0 / (double)(c_known + 1));
        ax.c = c_known;
        ax.coeff = w.flags.p + 3;
        ax.cur = w.fac; ax.old = w.old; ax.out = w.acc; ax.snap = snaps[it & 1];
      }
      if (o.normalization && !all_orth) {
        double *gp[kMaxOrder], *fp[kMaxOrder];
        for (int k = 0; k < N; ++k) { gp[k] = w.G(k); fp[k] = w.F(k); }
        normalize(ms, N, w.dims, R, w.ldr, wf, gp, fp, w.nscr, &ex, step_now ? &ax : nullptr);
      } else {
        export_words(ms, ex);
        if (step_now) axpy_accel(ms, w.fsz, ax.step_c, ax.step_c1, ax.c, ax.coeff, w.old, w.fac, w.acc, ax.snap);
      }
      bool tried = false, snapped = false;
      if (!cap && o.acceleration) {
        if (it > 1) {
          tried = true;
          snapped = true;
          // main stream first (the paired pass needs only A_hat), then the side stream:
          // the accelerated factors' Grams and, for the paired pass, mode 0 of the next
          // sweep for either outcome
          fork();
          const int cm = pair ? 0 : last;
          const double *mfit = w.M;
          if (pair) {
            left_node_pair(w, w.fac, w.acc);
            mfit = w.M2;
          } else {
            mttkrp_standard(w, w.acc, last, w.M);
          }
          // the sweep's cost (Normalize and the step leave M, A_last, V_last and G_last as
          // they are when the weight factor is not the last; the decision reads it after
          // the join)
          if (side_cost)
            cost_gram(ss, w.dims[last] * w.ldr, w.M, w.F(last), R, sel[last].V, w.G(last), w.scal.p, w.scal.p + 1, 0);
          if (pair) prep_mode(w, w.grams, 0, tree, slot(w, 0), o.max_inner, ss);
          for (int k = 0; k < N; ++k)
            gram(ss, w.dims[k], R, w.ldr, w.Fb(w.acc, k), w.agr.p + (int64_t)k * R * R, w.gpart_side);
          if (pair) prep_mode(w, w.agr, 0, tree, slot(w, N), o.max_inner, ss);
          // accel_cost_function (CpdDimTree.hpp:174-181): its fit term from the last mode's
          // MTTKRP, or from mode 0 of the paired left-node pass
          int list[kMaxOrder];
          int nl = v_order(N, cm, false, list);
          const double *gp[kMaxOrder];
          for (int j = 0; j < nl; ++j) gp[j] = w.agr.p + (int64_t)list[j] * R * R;
          hadamard(ss, R, nl, gp, w.Vacc);
          join();                    // the side stream has finished with the Grams, V_hat and slots
          cost_gram(ms, w.dims[cm] * w.ldr, mfit, w.Fb(w.acc, cm), R, w.Vacc, w.agr.p + (int64_t)cm * R * R,
                    w.scal.p, w.scal.p + 2, 1);
          AccelCopy cp{};
          cp.src[0] = w.acc; cp.dst[0] = w.fac; cp.n[0] = w.fsz;
          cp.src[1] = w.agr; cp.dst[1] = w.grams; cp.n[1] = (int64_t)N * R * R;
          if (pair) {                // the accelerated candidate's mode-0 set-up into slot 0
            const ModeIn a = slot(w, N), b = slot(w, 0);
            cp.src[2] = a.V; cp.dst[2] = b.V; cp.n[2] = (int64_t)R * R;
            cp.src[3] = a.betas; cp.dst[3] = b.betas; cp.n[3] = kMaxInner + 1;
            cp.prm_src = a.p; cp.prm_dst = b.p;
            const int64_t m0 = w.dims[0] * w.ldr;     // the mode-0 MTTKRP of the kept factors
            cp.pick_src[0] = w.M2.p; cp.pick_src[1] = w.M2.p + m0; cp.pick_dst = w.M; cp.pick_n = m0;
          }
          cp.fa_out = &rbdev->fa;
          cp.flags_out = rbdev->flags;
          cp.f_out = side_cost ? &rbdev->f : nullptr;
          if (!w.kev) { cp.seq_dst = &rbdev->seq; cp.seq = (unsigned)it; }
          accel_decide(ms, w.scal.p, o.threshold_error, w.flags, cp);
          if (pair) {
            // the next sweep's left node: the half of the pair the device kept
            w.left.p64 = nullptr;
            w.left.p32 = w.YL2.p;
            w.left.ld = 2 * w.rpf;
            w.left.sel = w.flags.p;
            w.left.sel_off = w.rpf;
            w.left_ready = true;
            sel[0] = slot(w, 0);
            prepped0 = true;
          }
        }
        else {
          HIPCHK(hipMemcpyAsync(w.old, w.fac, w.fsz * sizeof(double), hipMemcpyDeviceToDevice, ms));
        }
      }
      if (!snapped) HIPCHK(hipMemcpyAsync(snaps[it & 1], w.fac, w.fsz * sizeof(double), hipMemcpyDeviceToDevice, ms));
      if (w.kev) HIPCHK(hipEventRecord(evs.e[it & 1], ms));   // timed runs read their pass events after it
      tried_at[it & 1] = tried;
    };

    auto process = [&](int it) {
      if (w.kev) {
        HIPCHK(hipEventSynchronize(evs.e[it & 1]));
      } else {
        // the sweep's last readback kernel publishes `it` after a system-scope fence; the
        // stream is queried now and then so that a failed launch ends the wait
        const volatile unsigned *sq = &rbs[it & 1]->seq;
        for (unsigned spin = 1; *sq != (unsigned)it; ++spin) {
          if ((spin & 255) == 0) {
            const hipError_t q = hipStreamQuery(ms);
            if (q == hipSuccess && *sq != (unsigned)it) {
              h->err = "sweep readback missing after the stream completed";
              throw Fail{PTN_ERR_HIP};
            }
            if (q != hipSuccess && q != hipErrorNotReady) HIPCHK(q);
          }
        }
        std::atomic_thread_fence(std::memory_order_acquire);
      }
      const Readback *rb = rbs[it & 1];
      ptn_sweep_log lg{};
      r.ao_iter = it;
      r.f_value = rb->f;
      r.rel_cost = r.f_value / std::sqrt(r.frob_sq);
      lg.f_value = r.f_value;
      lg.rel_cost = r.rel_cost;
      for (int n = 0; n < N; ++n) {
        lg.nnls_iters[n] = rb->hp[n].iters + rb->hp[n].warm_steps;
        lg.warm_steps[n] = rb->hp[n].warm_steps;
        lg.zero_col_fallback[n] = rb->hp[n].fallback;
        lg.lipschitz[n] = rb->hp[n].L0;
        lg.mu[n] = rb->hp[n].mu0;
        lg.warm_rows[n] = rb->hp[n].warm_rows;
      }
      for (int q = 0; q < 2 && w.kev; ++q)
        if (w.kev_used[it & 1][q]) {              // the sweep's events are complete (its end event is)
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, w.kev[it & 1][2 * q], w.kev[it & 1][2 * q + 1]));
          lg.t_pass_ms[q] = ms;
          w.kev_used[it & 1][q] = false;
        }
      const bool stop = r.rel_cost < o.threshold_error || it >= (int)o.max_iter;
      if (tried_at[it & 1] && !stop) {
        lg.accel_tried = 1;
        lg.f_accel = rb->fa;
        if (rb->flags[0]) {
          r.f_value = rb->fa;
          lg.accel_accepted = 1;
        }
        r.accel_fail = rb->flags[2];
        r.accel_coeff = rb->flags[3];
        c_known = rb->flags[3];
      }
      lg.t_end = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (log && it - 1 < log_cap) log[it - 1] = lg;
      return stop;
    };

    t0 = std::chrono::steady_clock::now();
    int final_it = 1;
    enqueue(1);
    for (int it = 1;; ++it) {
      if (it < (int)o.max_iter) enqueue(it + 1);
      if (process(it)) {
        final_it = it;
        break;
      }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipStreamSynchronize(h->side));
    HIPCHK(hipStreamSynchronize(h->aux));
    r.sweep_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    HIPCHK(hipGetLastError());
    download_factors(w, snaps[final_it & 1], out_packed);
    *res = r;
  });
}

ptn_status ptn_mttkrp(ptn_handle h, int R, const double *factors_packed, int mode, int arith, double *M_out) {
  return run(h, [&] {
    check_arith(h, arith);
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(mode >= 0 && mode < h->N, PTN_ERR_INVALID, "bad mode");
    REQUIRE(factors_packed && M_out, PTN_ERR_INVALID, "null argument");
    Work w(h, R, arith, false);
    upload_factors(w, factors_packed, w.fac);
    mttkrp_standard(w, w.fac, mode, w.M);
    DevBuf<double> cm(h, w.dims[mode] * R);
    rows_to_colmajor(h->stream, w.dims[mode], R, w.ldr, w.M, cm.p);
    HIPCHK(hipMemcpyAsync(M_out, cm.p, w.dims[mode] * R * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
  });
}

ptn_status ptn_mttkrp_dimtree(ptn_handle h, int R, const double *factors_packed, int arith, double *M_packed_out) {
  return run(h, [&] {
    check_arith(h, arith);
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(h->N >= 3 && h->N <= kMaxOrder, PTN_ERR_UNSUPPORTED, "dimension tree built for N = 3 .. 8");
    REQUIRE(factors_packed && M_packed_out, PTN_ERR_INVALID, "null argument");
    Work w(h, R, arith, true);
    upload_factors(w, factors_packed, w.fac);
    int64_t o = 0;
    for (int n = 0; n < w.N; ++n) {
      mttkrp_tree(w, w.fac, n, w.M);
      DevBuf<double> cm(h, w.dims[n] * R);
      rows_to_colmajor(h->stream, w.dims[n], R, w.ldr, w.M, cm.p);
      HIPCHK(hipMemcpyAsync(M_packed_out + o, cm.p, w.dims[n] * R * sizeof(double), hipMemcpyDeviceToHost,
                            h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      o += w.dims[n] * R;
    }
    HIPCHK(hipGetLastError());
  });
}

ptn_status ptn_sym_extremes(ptn_handle h, int R, const double *V, double *L, double *mu) {
  return run(h, [&] {
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(V && L && mu, PTN_ERR_INVALID, "null argument");
    DevBuf<double> dV(h, (size_t)R * R);
    DevBuf<NnlsParams> p(h, 1);
    HIPCHK(hipMemcpyAsync(dV.p, V, (size_t)R * R * sizeof(double), hipMemcpyHostToDevice, h->stream));
    eig_extremes(h->stream, R, dV.p, p.p);
    NnlsParams hp;
    HIPCHK(hipMemcpyAsync(&hp, p.p, sizeof(hp), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    *L = hp.L0;
    *mu = hp.mu0;
  });
}

ptn_status ptn_nesterov_mnls(ptn_handle h, int64_t I, int R, const double *Z, const double *W, double delta_1,
                             double delta_2, int max_inner, double *A, int *iters, double *L, double *mu) {
  return run(h, [&] {
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(I >= 1 && I < (1ll << 31), PTN_ERR_INVALID, "bad row count");
    REQUIRE(max_inner >= 0 && max_inner <= kMaxInner, PTN_ERR_UNSUPPORTED, "max_inner must be in [0, 255]");
    REQUIRE(Z && W && A, PTN_ERR_INVALID, "null argument");
    const int ldr = round_up16(R);
    DevBuf<double> dZ(h, (size_t)R * R), tmp(h, I * R), dW(h, I * ldr), dA(h, I * ldr), dB(h, I * ldr);
    DevBuf<double> hist(h, (size_t)(max_inner + 1) * I * ldr), betas(h, kMaxInner + 1);
    DevBuf<double> colsum(h, kMaxRank);
    DevBuf<NnlsParams> p(h, 1);
    DevBuf<uint64_t> mask(h, nnls_mask_words(I));
    std::vector<double> zr((size_t)R * R);
    for (int a = 0; a < R; ++a)
      for (int b = 0; b < R; ++b) zr[a * R + b] = Z[a + (int64_t)R * b];
    HIPCHK(hipMemcpyAsync(dZ.p, zr.data(), zr.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(tmp.p, W, I * R * sizeof(double), hipMemcpyHostToDevice, h->stream));
    colmajor_to_rows(h->stream, I, R, ldr, tmp.p, dW.p);
    HIPCHK(hipMemcpyAsync(tmp.p, A, I * R * sizeof(double), hipMemcpyHostToDevice, h->stream));
    colmajor_to_rows(h->stream, I, R, ldr, tmp.p, dA.p);
    eig_extremes(h->stream, R, dZ.p, p.p);
    nnls_prep(h->stream, p.p, max_inner, betas.p);
    nnls_iterate(h->stream, I, R, ldr, dZ.p, p.p, betas.p, dW.p, dA.p, nullptr, delta_1, delta_2, max_inner, mask.p,
                 hist.p, colsum.p);
    nnls_select(h->stream, I, R, ldr, hist.p, mask.p, max_inner, dB.p, colsum.p, p.p);
    rows_to_colmajor(h->stream, I, R, ldr, dB.p, tmp.p);
    HIPCHK(hipMemcpyAsync(A, tmp.p, I * R * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    NnlsParams hp;
    HIPCHK(hipMemcpyAsync(&hp, p.p, sizeof(hp), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    if (iters) *iters = hp.iters;
    if (L) *L = hp.L0;
    if (mu) *mu = hp.mu0;
  });
}

}  // extern "C"

namespace {
// run() without the std::function (no host heap use on the no-allocation path).
template <typename F>
ptn_status run_direct(ptn_handle h, F &&fn) {
  if (!h) return PTN_ERR_INVALID;
  try {
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) {
      h->err = hipGetErrorString(e);
      return PTN_ERR_HIP;
    }
    fn();
    return PTN_OK;
  } catch (const Fail &f) {
    return f.code;
  } catch (...) {
    h->err = "unexpected exception";
    return PTN_ERR_HIP;
  }
}

int64_t align256(int64_t bytes) { return (bytes + 255) / 256 * 256; }
}  // namespace

extern "C" {

ptn_status ptn_nesterov_mnls_noalloc(ptn_handle h, int64_t I, int R, const double *Z, const double *W,
                                     double delta_1, double delta_2, int max_inner, double *A, int *iters, double *L,
                                     double *mu) {
  return run_direct(h, [&] {
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(I >= 1 && I < (1ll << 31), PTN_ERR_INVALID, "bad row count");
    REQUIRE(max_inner >= 0 && max_inner <= kMaxInner, PTN_ERR_UNSUPPORTED, "max_inner must be in [0, 255]");
    REQUIRE(Z && W && A, PTN_ERR_INVALID, "null argument");
    REQUIRE(h->device >= 0 && h->device < kNoallocMaxDevices && noalloc_ptr[h->device], PTN_ERR_STATE,
            "no-allocation scratch not set up for this device");
    const int ldr = round_up16(R);
    // scratch layout, 256-byte aligned pieces: Z rows, W rows, A rows, result rows, the
    // iterates, betas, column sums, NnlsParams, satisfaction masks
    const int64_t rows = I * ldr * (int64_t)sizeof(double);
    const int64_t o_z = 0, o_w = o_z + align256((int64_t)R * R * sizeof(double)), o_a = o_w + align256(rows),
                  o_b = o_a + align256(rows), o_hist = o_b + align256(rows),
                  o_beta = o_hist + align256((int64_t)(max_inner + 1) * rows),
                  o_cs = o_beta + align256((kMaxInner + 1) * (int64_t)sizeof(double)),
                  o_prm = o_cs + align256(kMaxRank * (int64_t)sizeof(double)),
                  o_mask = o_prm + align256((int64_t)sizeof(NnlsParams)),
                  end = o_mask + align256(nnls_mask_words(I) * (int64_t)sizeof(uint64_t));
    if (end > (int64_t)PTN_NOALLOC_SCRATCH_BYTES) {
      h->err = "ptn_nesterov_mnls_noalloc: I = " + std::to_string(I) + ", R = " + std::to_string(R) +
               ", max_inner = " + std::to_string(max_inner) + " needs " + std::to_string(end) +
               " bytes of scratch, above the PTN_NOALLOC_SCRATCH_BYTES limit of " +
               std::to_string((long long)PTN_NOALLOC_SCRATCH_BYTES) + "; use ptn_nesterov_mnls";
      throw Fail{PTN_ERR_INVALID};
    }
    unsigned char *base = (unsigned char *)noalloc_ptr[h->device];
    double *dZ = (double *)(base + o_z), *dW = (double *)(base + o_w), *dA = (double *)(base + o_a),
           *dB = (double *)(base + o_b), *hist = (double *)(base + o_hist), *betas = (double *)(base + o_beta),
           *colsum = (double *)(base + o_cs);
    NnlsParams *prm = (NnlsParams *)(base + o_prm);
    uint64_t *mask = (uint64_t *)(base + o_mask);
    hipStream_t s = h->stream;
    // the scratch keeps the previous call's contents: betas, column sums, NnlsParams and
    // masks start at zero as in ptn_nesterov_mnls (every other piece is written before it
    // is read)
    fill_zero(s, base + o_beta, (end - o_beta) / 8);
    colmajor_to_rows(s, R, R, R, Z, dZ);            // row-major Z (Z[a + R b] -> dZ[a R + b])
    colmajor_to_rows(s, I, R, ldr, W, dW);
    colmajor_to_rows(s, I, R, ldr, A, dA);
    eig_extremes(s, R, dZ, prm);
    nnls_prep(s, prm, max_inner, betas);
    nnls_iterate(s, I, R, ldr, dZ, prm, betas, dW, dA, nullptr, delta_1, delta_2, max_inner, mask, hist, colsum);
    nnls_select(s, I, R, ldr, hist, mask, max_inner, dB, colsum, prm);
    rows_to_colmajor(s, I, R, ldr, dB, A);
    HIPCHK(hipGetLastError());
    if (iters || L || mu) {
      NnlsParams hp;
      HIPCHK(hipMemcpyAsync(&hp, prm, sizeof(hp), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (iters) *iters = hp.iters;
      if (L) *L = hp.L0;
      if (mu) *mu = hp.mu0;
    } else {
      HIPCHK(hipStreamSynchronize(s));
    }
  });
}

ptn_status ptn_normalize(ptn_handle h, int N, const int64_t *dims, int R, int weight_factor, double *grams_packed,
                         double *factors_packed) {
  return run(h, [&] {
    check_dims(h, N, dims);
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(weight_factor >= 0 && weight_factor < N, PTN_ERR_INVALID, "bad weight factor");
    REQUIRE(grams_packed && factors_packed, PTN_ERR_INVALID, "null argument");
    const int ldr = round_up16(R);
    int64_t off[kMaxOrder], o = 0, nf = 0;
    for (int k = 0; k < N; ++k) { off[k] = o; o += dims[k] * ldr; nf += dims[k] * R; }
    DevBuf<double> fac(h, o), tmp(h, nf), G(h, (size_t)N * R * R), scr(h, 2 * kMaxOrder * kMaxRank);
    std::vector<double> gr((size_t)N * R * R);
    for (int k = 0; k < N; ++k)
      for (int a = 0; a < R; ++a)
        for (int b = 0; b < R; ++b) gr[(size_t)k * R * R + a * R + b] = grams_packed[(size_t)k * R * R + a + R * b];
    HIPCHK(hipMemcpyAsync(G.p, gr.data(), gr.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(tmp.p, factors_packed, nf * sizeof(double), hipMemcpyHostToDevice, h->stream));
    int64_t oc = 0;
    double *gp[kMaxOrder], *fp[kMaxOrder];
    for (int k = 0; k < N; ++k) {
      colmajor_to_rows(h->stream, dims[k], R, ldr, tmp.p + oc, fac.p + off[k]);
      oc += dims[k] * R;
      gp[k] = G.p + (int64_t)k * R * R;
      fp[k] = fac.p + off[k];
    }
    normalize(h->stream, N, dims, R, ldr, weight_factor, gp, fp, scr.p);
    oc = 0;
    for (int k = 0; k < N; ++k) {
      rows_to_colmajor(h->stream, dims[k], R, ldr, fac.p + off[k], tmp.p + oc);
      oc += dims[k] * R;
    }
    HIPCHK(hipMemcpyAsync(factors_packed, tmp.p, nf * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(gr.data(), G.p, gr.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    for (int k = 0; k < N; ++k)
      for (int a = 0; a < R; ++a)
        for (int b = 0; b < R; ++b) grams_packed[(size_t)k * R * R + a + R * b] = gr[(size_t)k * R * R + a * R + b];
  });
}

ptn_status ptn_cost_explicit(ptn_handle h, int R, const double *factors_packed, double *f) {
  return run(h, [&] {
    REQUIRE(h->N > 0, PTN_ERR_STATE, "no tensor set on the handle");
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(factors_packed && f, PTN_ERR_INVALID, "null argument");
    Work w(h, R, h->x64 ? PTN_ARITH_F64 : 3, false);
    upload_factors(w, factors_packed, w.fac);
    *f = cost_explicit(w, w.fac);
  });
}

ptn_status ptn_time_contraction(ptn_handle h, int kernel, int R, const double *factors_packed, int arith, int iters,
                                double *avg_ms, double *bytes_out) {
  return run(h, [&] {
    check_arith(h, arith);
    REQUIRE(R >= 1 && R <= PTN_MAX_RANK, PTN_ERR_UNSUPPORTED, "rank must be in [1, 64]");
    REQUIRE(kernel >= 0 && kernel <= 2 && iters >= 1 && avg_ms, PTN_ERR_INVALID, "bad argument");
    REQUIRE(h->N == 3 || kernel == 2 || (kernel == 0 && h->N == 4), PTN_ERR_UNSUPPORTED,
            "node timing built for N = 3 (and the order-4 left node)");
    Work w(h, R, arith, true);
    upload_factors(w, factors_packed, w.fac);
    const int64_t *d = w.dims;
    const double sx = arith == PTN_ARITH_F64 ? 8.0 : 4.0;
    // operands packed once; the timed launches are the contraction kernel alone
    const bool fused = kernel == 1 && w.mfma32 && w.N == 3;
    const bool n4 = w.N == 4;
    // order 4: the fused GEMM + TTV node X_(0,1) (A_3 (.) A_2), split over k and reduced
    const int S4 = n4 ? pick_splits_cm(w, d[0] * d[1], d[2] * d[3]) : 1;
    if (w.mfma32) {
      if (kernel == 0) w.pack(n4 ? kr_of(w, w.fac, {2, 3}) : kr_of(w, w.fac, {2}), n4 ? d[2] * d[3] : d[2], 0);
      else if (kernel == 1) w.pack(kr_of(w, w.fac, {0}), d[0], 1);
    }
    auto launch = [&] {
      if (kernel == 0 && n4) {
        const int64_t Mr = d[0] * d[1], K = d[2] * d[3];
        if (w.mfma32) {
          REQUIRE(S4 * Mr * w.rpf <= w.part32_cap, PTN_ERR_INVALID, "partial buffer too small");
          gemm_cm_f32(h->stream, w.gmode(), w.x32(), Mr, K, w.rpf, S4, w.whi, w.wlo, w.wf, w.part32);
          reduce_splits_f32(h->stream, w.part32, S4, Mr * w.rpf, w.YL32);
        } else {
          node_cm(w, kr_of(w, w.fac, {2, 3}), Mr, K, S4, w.YL.p, w.YL32.p);
        }
      } else if (kernel == 0) {
        if (w.mfma32) gemm_cm_f32(h->stream, w.gmode(), w.x32(), d[0] * d[1], d[2], w.rpf, 1, w.whi, w.wlo, w.wf,
                                  w.YL32);
        else node_cm(w, kr_of(w, w.fac, {2}), d[0] * d[1], d[2], 1, w.YL.p, w.YL32.p);
      } else if (kernel == 1) {
        if (fused) gemm_ck_f32(h->stream, w.gmode(), w.x32(), d[0], d[1] * d[2], w.rpf, w.whi, w.wlo, w.wf, nullptr,
                               d[1], w.F(1), w.ldr, w.part);
        else if (w.mfma32) gemm_ck_f32(h->stream, w.gmode(), w.x32(), d[0], w.total / d[0], w.rpf, w.whi, w.wlo,
                                       w.wf, w.T32, 0, nullptr, w.ldr, nullptr);
        else node_ck(w, kr_of(w, w.fac, {0}), d[0], w.total / d[0]);
      } else {
        mttkrp_standard(w, w.fac, 0, w.M);
      }
    };
    launch();
    HIPCHK(hipStreamSynchronize(h->stream));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, h->stream));
    for (int i = 0; i < iters; ++i) launch();
    HIPCHK(hipEventRecord(e1, h->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    *avg_ms = ms / iters;
    if (bytes_out) {
      // algorithmic bytes: tensor once + output + factor reads
      double b = sx * (double)w.total;
      const double sy = w.mfma32 ? 4.0 : 8.0;   // node element size
      if (kernel == 0 && n4) b += sy * R * (double)(d[0] * d[1]) + 8.0 * R * (double)(d[2] + d[3]);
      else if (kernel == 0) b += sy * R * (double)(d[0] * d[1]) + 8.0 * R * d[2];
      else if (fused) b += 8.0 * R * d[2] + 8.0 * R * (double)(d[0] + d[1]);
      else if (kernel == 1) b += sy * R * (double)(w.total / d[0]) + 8.0 * R * d[0];
      else b += 8.0 * R * d[0] + 8.0 * R * (double)(d[1] + d[2]);
      *bytes_out = b;
    }
  });
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// batched independent order-3 problems
// ------------------------------------------------------------------------------------
namespace {

constexpr int kBatchBR = 16;        // factor rows of a batched problem hold 16 doubles
constexpr int kBatchMaxDim = 128;
constexpr int kBatchMaxRank = 16;

static_assert(sizeof(BatchLog) == sizeof(ptn_batch_log), "ptn_batch_log layout");

void batch_layout(ptn_handle h, int nprob, const int *dims, const int *ranks) {
  REQUIRE(nprob >= 1 && nprob <= (1 << 20), PTN_ERR_INVALID, "nprob must be in [1, 2^20]");
  REQUIRE(dims && ranks, PTN_ERR_INVALID, "dims or ranks is null");
  std::unique_ptr<BatchSet> b(new BatchSet());
  b->nprob = nprob;
  b->desc.resize(nprob);
  int64_t xo = 0, fo = 0;
  for (int p = 0; p < nprob; ++p) {
    BatchDesc &d = b->desc[p];
    std::memset(&d, 0, sizeof(d));
    for (int k = 0; k < 3; ++k) {
      REQUIRE(dims[3 * p + k] >= 1 && dims[3 * p + k] <= kBatchMaxDim, PTN_ERR_UNSUPPORTED,
              "batched problems take mode sizes in [1, 128]");
      d.I[k] = dims[3 * p + k];
    }
    REQUIRE(ranks[p] >= 1 && ranks[p] <= kBatchMaxRank, PTN_ERR_UNSUPPORTED, "batched problems take ranks in [1, 16]");
    d.R = ranks[p];
    d.rowoff[0] = 0;
    d.rowoff[1] = d.I[0];
    d.rowoff[2] = d.I[0] + d.I[1];
    d.x_off = xo;
    xo += ((int64_t)d.I[0] * d.I[1] * d.I[2] + 63) / 64 * 64;
    d.f_off = fo;
    fo += d.I[0] + d.I[1] + d.I[2];
    d.g_off = (int64_t)p * 3 * kBatchBR * kBatchBR;
  }
  b->nx = xo;
  b->nrows = fo;
  // three copies: mode 0 fastest (the tensors as given), mode 1 fastest, mode 2 fastest
  HIPCHK(hipMalloc((void **)&b->x32, 3 * xo * sizeof(float)));
  HIPCHK(hipMemsetAsync(b->x32, 0, 3 * xo * sizeof(float), h->stream));
  delete h->batch;
  h->batch = b.release();
}

int64_t batch_elems(const BatchDesc &d) { return (int64_t)d.I[0] * d.I[1] * d.I[2]; }

// the per-mode copies every sweep's MTTKRPs read (each mode's rows contiguous)
void batch_copies(ptn_handle h) {
  BatchSet &b = *h->batch;
  for (const BatchDesc &d : b.desc)
    batch_permute(h->stream, b.x32 + d.x_off, d.I[0], d.I[1], d.I[2], b.x32 + b.nx + d.x_off,
                  b.x32 + 2 * b.nx + d.x_off);
}

struct GraphExec {                       // a captured and instantiated graph, destroyed with the scope
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  ~GraphExec() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
  }
};

struct TimingEvents {
  hipEvent_t e[2] = {nullptr, nullptr};
  explicit TimingEvents(ptn_handle h) {
    for (hipEvent_t &x : e) HIPCHK(hipEventCreate(&x));
  }
  ~TimingEvents() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
  double ms(ptn_handle h) {
    HIPCHK(hipEventSynchronize(e[1]));
    float v = 0.f;
    HIPCHK(hipEventElapsedTime(&v, e[0], e[1]));
    return v;
  }
};

}  // namespace

extern "C" {

ptn_status ptn_batch_set(ptn_handle h, int nprob, const int *dims, const int *ranks, const double *X) {
  return run(h, [&] {
    REQUIRE(X != nullptr, PTN_ERR_INVALID, "X is null");
    batch_layout(h, nprob, dims, ranks);
    BatchSet &b = *h->batch;
    std::vector<float> xs(b.nx, 0.f);
    int64_t src = 0;
    for (const BatchDesc &d : b.desc) {
      const int64_t n = batch_elems(d);
      for (int64_t e = 0; e < n; ++e) xs[d.x_off + e] = (float)X[src + e];
      src += n;
    }
    HIPCHK(hipMemcpyAsync(b.x32, xs.data(), b.nx * sizeof(float), hipMemcpyHostToDevice, h->stream));
    batch_copies(h);
    HIPCHK(hipStreamSynchronize(h->stream));
  });
}

ptn_status ptn_batch_generate(ptn_handle h, int nprob, const int *dims, const int *ranks, uint64_t seed,
                              double snr_db) {
  return run(h, [&] {
    batch_layout(h, nprob, dims, ranks);
    BatchSet &b = *h->batch;
    int64_t maxn = 1, maxf = 1;
    for (const BatchDesc &d : b.desc) {
      maxn = std::max(maxn, batch_elems(d));
      maxf = std::max(maxf, (int64_t)(d.I[0] + d.I[1] + d.I[2]) * d.R);
    }
    DevBuf<double> x64(h, maxn), fac(h, maxf);
    for (int p = 0; p < nprob; ++p) {
      const BatchDesc &d = b.desc[p];
      const int64_t dd[3] = {d.I[0], d.I[1], d.I[2]};
      const int64_t n = batch_elems(d);
      const uint64_t sp = seed + (uint64_t)p;
      gen_factors(h->stream, 3, dd, d.R, sp, 0, fac.p);
      cp_full(h->stream, 3, dd, d.R, fac.p, x64.p, nullptr);
      if (!std::isnan(snr_db)) {
        square_norm_partials(h->stream, n, x64.p, nullptr, h->part, kNormParts);
        sum_partials(h->stream, h->part, kNormParts, h->scal + 1);
        gen_noise_norm(h->stream, n, sp, h->part, kNormParts);
        sum_partials(h->stream, h->part, kNormParts, h->scal + 2);
        noise_scale(h->stream, h->scal + 1, h->scal + 2, snr_db, h->scal + 3);
        add_noise(h->stream, n, sp, h->scal + 3, x64.p, nullptr);
      }
      cast_f64_to_f32(h->stream, n, x64.p, b.x32 + d.x_off);
    }
    batch_copies(h);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
  });
}

ptn_status ptn_batch_init_factors(ptn_handle h, uint64_t seed, double *packed_out) {
  return run(h, [&] {
    REQUIRE(h->batch, PTN_ERR_STATE, "no batch set on the handle");
    REQUIRE(packed_out, PTN_ERR_INVALID, "output is null");
    const BatchSet &b = *h->batch;
    int64_t tot = 0;
    for (const BatchDesc &d : b.desc) tot += (int64_t)(d.I[0] + d.I[1] + d.I[2]) * d.R;
    DevBuf<double> fac(h, tot);
    int64_t off = 0;
    for (int p = 0; p < b.nprob; ++p) {
      const BatchDesc &d = b.desc[p];
      const int64_t dd[3] = {d.I[0], d.I[1], d.I[2]};
      gen_factors(h->stream, 3, dd, d.R, seed + (uint64_t)p, 0, fac.p + off);
      off += (int64_t)(d.I[0] + d.I[1] + d.I[2]) * d.R;
    }
    HIPCHK(hipMemcpyAsync(packed_out, fac.p, tot * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
  });
}

ptn_status ptn_batch_get_tensor(ptn_handle h, int p, double *X_out) {
  return run(h, [&] {
    REQUIRE(h->batch, PTN_ERR_STATE, "no batch set on the handle");
    REQUIRE(p >= 0 && p < h->batch->nprob, PTN_ERR_INVALID, "no such problem");
    REQUIRE(X_out, PTN_ERR_INVALID, "output is null");
    const BatchDesc &d = h->batch->desc[p];
    const int64_t n = batch_elems(d);
    std::vector<float> xs(n);
    HIPCHK(hipMemcpyAsync(xs.data(), h->batch->x32 + d.x_off, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int64_t e = 0; e < n; ++e) X_out[e] = (double)xs[e];
  });
}

ptn_status ptn_cpd_batch(ptn_handle h, const double *init_packed, const ptn_options *opt, double *out_packed,
                         ptn_result *res, ptn_batch_log *log, int block, int grid, ptn_batch_report *rep) {
  return run(h, [&] {
    REQUIRE(h->batch, PTN_ERR_STATE, "no batch set on the handle");
    REQUIRE(init_packed && opt && out_packed && res, PTN_ERR_INVALID, "null argument");
    const ptn_options &o = *opt;
    REQUIRE(o.max_iter >= 1 && o.max_iter <= 1000000, PTN_ERR_INVALID, "max_iter must be in [1, 1e6]");
    REQUIRE(o.max_inner >= 0 && o.max_inner <= kMaxInner, PTN_ERR_UNSUPPORTED, "max_inner must be in [0, 255]");
    REQUIRE(o.accel_coeff >= 1, PTN_ERR_INVALID, "accel_coeff must be >= 1");
    REQUIRE(o.accel_fail >= 0, PTN_ERR_INVALID, "accel_fail must be >= 0");
    for (int k = 0; k < 3; ++k)
      REQUIRE(o.constraints[k] == PTN_NONNEGATIVITY, PTN_ERR_UNSUPPORTED, "only the nonnegativity constraint is built");
    REQUIRE((block == 0 && grid == 0) || ((block == 256 || block == 512) && grid >= 1 && grid <= 8192),
            PTN_ERR_INVALID, "block/grid: both 0 (autotune) or block 256 / 512 with grid in [1, 8192]");
    REQUIRE(o.arith == PTN_ARITH_F64 || o.arith == PTN_ARITH_F32, PTN_ERR_UNSUPPORTED,
            "batch MTTKRP arithmetic: PTN_ARITH_F64 (FP64 matrix cores) or PTN_ARITH_F32 (FP32 matrix cores)");
    const BatchSet &b = *h->batch;
    const int P = b.nprob;
    const int64_t nr = b.nrows * kBatchBR;
    hipStream_t s = h->stream;

    // factor rows (16 doubles, zero beyond R) from the packed column-major factors
    std::vector<double> rows(nr, 0.0);
    int64_t src = 0;
    for (const BatchDesc &d : b.desc) {
      for (int k = 0; k < 3; ++k) {
        const int I = d.I[k];
        for (int r = 0; r < d.R; ++r)
          for (int i = 0; i < I; ++i) rows[(d.f_off + d.rowoff[k] + i) * kBatchBR + r] = init_packed[src + i + (int64_t)I * r];
        src += (int64_t)I * d.R;
      }
    }
    std::vector<BatchProb> st0(P);
    for (BatchProb &q : st0) {
      std::memset(&q, 0, sizeof(q));
      q.accel_coeff = o.accel_coeff;
      q.accel_fail = o.accel_fail;
    }
    // Work of a problem: its tensor elements plus kRowWork per factor row.  One block's time on
    // a problem (wall-clock stamps per problem over 100 sweeps of the 512-problem batch) is
    // ~0.2 ns per element and MTTKRP pass whatever the rank (the matrix-core tile is 16 columns
    // wide), plus ~0.35 us per factor row (Nesterov steps, Grams, Normalize): a rank-8
    // 122 x 128 x 104 problem took as long as a rank-15 123 x 121 x 126 one cut in two, so the
    // work is not weighted by the rank.  A problem with more than kSplitWork times the batch's
    // mean work has its MTTKRP sums cut in V = ceil(work / (kSplitWork mean)) pieces (at most
    // 8), so that up to V blocks share it (BatchDesc::V); the pieces, and with them every
    // result, depend on the batch only.  The other members repeat the solves, so splitting
    // pays only for the largest problems: on the 512-problem batch (512 x 256 launch) a factor
    // of 3 gave 1.21 ms per sweep (FP32 matrix cores) and 1.29 ms (FP64 ones), against 1.39 and
    // 1.55 ms with the rank-weighted work and a factor of 4.
    constexpr double kSplitWork = 3.0;
    constexpr double kRowWork = 540.0;
    std::vector<BatchDesc> dv(b.desc);
    auto work = [&](int p) {
      const BatchDesc &d = b.desc[p];
      return (double)batch_elems(d) + kRowWork * (d.I[0] + d.I[1] + d.I[2]);
    };
    double mean = 0.0;
    for (int p = 0; p < P; ++p) mean += work(p) / P;
    int64_t hrows = 0, hgr = 0, mxs = 0;
    for (int p = 0; p < P; ++p) {
      BatchDesc &d = dv[p];
      d.V = std::max(1, std::min(kBatchMaxMembers, (int)std::ceil(work(p) / (kSplitWork * mean))));
      d.G = 1;
      d.h_off = hrows;
      d.hg_off = hgr;
      d.mx_off = mxs;
      hrows += (int64_t)(d.V - 1) * (d.I[0] + d.I[1] + d.I[2]);
      hgr += (int64_t)(d.V - 1) * 3 * kBatchBR * kBatchBR;
      if (d.V > 1) mxs += (int64_t)2 * d.V * kBatchMaxDim * kBatchBR;
    }
    // Queue of one launch shape: each problem runs on G = min(V, grid) blocks (members next
    // to each other in the queue), problems ordered by the work of one member, largest first
    std::vector<int> order;
    auto shape_queue = [&](int grid_blocks) {
      std::vector<int> pr(P);
      std::iota(pr.begin(), pr.end(), 0);
      for (int p = 0; p < P; ++p) dv[p].G = std::min(dv[p].V, grid_blocks);
      std::stable_sort(pr.begin(), pr.end(), [&](int x, int y) { return work(x) / dv[x].G > work(y) / dv[y].G; });
      order.clear();
      for (int p : pr)
        for (int g = 0; g < dv[p].G; ++g) order.push_back(8 * p + g);
    };
    // (sweep + 1)^(1 / coeff) for every sweep and every coefficient the bookkeeping can reach
    const int stride = o.accel_coeff + o.max_iter / 5 + 3;
    std::vector<double> pw((int64_t)(o.max_iter + 1) * stride, 0.0);
    for (int sw = 0; sw <= o.max_iter; ++sw)
      for (int c = 1; c < stride; ++c) pw[(int64_t)sw * stride + c] = std::pow((double)(sw + 1), 1.0 / (double)c);

    DevBuf<BatchDesc> desc(h, P);
    DevBuf<BatchProb> st(h, P), st_snap(h, P);
    DevBuf<double> fac(h, nr), old(h, nr), acc(h, nr), grams(h, (size_t)P * 3 * kBatchBR * kBatchBR);
    DevBuf<double> fac_snap(h, nr), old_snap(h, nr), grams_snap(h, grams.n), powt(h, pw.size());
    DevBuf<double> hfac(h, hrows * kBatchBR), hold(h, hrows * kBatchBR), hacc(h, hrows * kBatchBR), hgrams(h, hgr),
        mx(h, mxs);
    DevBuf<int> ord(h, (size_t)P * kBatchMaxMembers), queue(h, o.max_iter + 2), ctr(h, 1), gctr(h, P), gerr(h, 1);
    int cu = 0;
    HIPCHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, h->device));
    // partial bands of 256-thread blocks: one slice per block of the largest grid launched
    DevBuf<double> pscr(h, (size_t)std::max(block == 0 ? 2 * cu : grid, 1) * kBatchPartDoubles256);
    DevBuf<BatchLog> lg(h, log ? (size_t)P * o.max_iter : 1);
    const int one = 1;
    HIPCHK(hipMemcpyAsync(st.p, st0.data(), P * sizeof(BatchProb), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(fac.p, rows.data(), nr * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(powt.p, pw.data(), pw.size() * sizeof(double), hipMemcpyHostToDevice, s));

    BatchArgs a{};
    a.nprob = P;
    a.desc = desc.p;
    a.st = st.p;
    a.X = b.x32;
    a.fac = fac.p;
    a.old = old.p;
    a.acc = acc.p;
    a.grams = grams.p;
    a.order = ord.p;
    a.queue = queue.p;
    a.sweep_ctr = ctr.p;
    a.max_iter = o.max_iter;
    a.thr = o.threshold_error;
    a.d1 = o.nesterov_delta_1;
    a.d2 = o.nesterov_delta_2;
    a.max_inner = o.max_inner;
    a.acceleration = o.acceleration;
    a.normalization = o.normalization;
    a.pow_table = powt.p;
    a.pow_stride = stride;
    a.log = nullptr;
    a.arith_f32 = o.arith == PTN_ARITH_F32 ? 1 : 0;
    a.xcopy = b.nx;
    a.hfac = hfac.p;
    a.hold = hold.p;
    a.hacc = hacc.p;
    a.hgrams = hgrams.p;
    a.mx = mx.p;
    a.gctr = gctr.p;
    a.err = gerr.p;
    a.pscratch = pscr.p;
    // the queue of a launch shape; BatchDesc::G follows it (hipMemcpyAsync from pageable
    // memory returns once the source has been read, so dv and order may change afterwards)
    auto set_shape = [&](int grid_blocks) {
      shape_queue(grid_blocks);
      HIPCHK(hipMemcpyAsync(desc.p, dv.data(), P * sizeof(BatchDesc), hipMemcpyHostToDevice, s));
      HIPCHK(hipMemcpyAsync(ord.p, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, s));
      HIPCHK(hipStreamSynchronize(s));
      a.nitems = (int)order.size();
    };
    set_shape(1);
    batch_init(s, a);                                   // ||X_p||^2 and the initial Grams
    // the host copy takes the norms, so that later uploads of dv keep them
    HIPCHK(hipMemcpyAsync(dv.data(), desc.p, P * sizeof(BatchDesc), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));

    // a run from the current state: sweep counters, group arrivals, the helpers' copies
    auto restart = [&] {
      HIPCHK(hipMemsetAsync(queue.p, 0, queue.n * sizeof(int), s));
      HIPCHK(hipMemsetAsync(gctr.p, 0, gctr.n * sizeof(int), s));
      HIPCHK(hipMemsetAsync(gerr.p, 0, sizeof(int), s));
      HIPCHK(hipMemcpyAsync(ctr.p, &one, sizeof(int), hipMemcpyHostToDevice, s));
      batch_sync_helpers(s, a);
    };
    double tune_s = 0.0;
    const bool tune = block == 0;
    if (tune) {
      // autotuner: a few launch shapes, each timed over the first two sweeps from the
      // same state (saved here and restored after every candidate)
      const auto t0 = std::chrono::steady_clock::now();
      HIPCHK(hipMemcpyAsync(st_snap.p, st.p, P * sizeof(BatchProb), hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(fac_snap.p, fac.p, nr * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(old_snap.p, old.p, nr * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(grams_snap.p, grams.p, grams.n * sizeof(double), hipMemcpyDeviceToDevice, s));
      auto restore = [&] {
        HIPCHK(hipMemcpyAsync(st.p, st_snap.p, P * sizeof(BatchProb), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(fac.p, fac_snap.p, nr * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(old.p, old_snap.p, nr * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(grams.p, grams_snap.p, grams.n * sizeof(double), hipMemcpyDeviceToDevice, s));
        restart();
      };
      const int cand[][2] = {{256, 2 * cu}, {256, cu}, {512, cu}, {512, cu / 2 > 0 ? cu / 2 : 1}};
      const int tsweeps = std::min(2, o.max_iter);
      double best = 1e300;
      TimingEvents ev(h);
      for (const auto &c : cand) {
        set_shape(c[1]);
        restore();
        HIPCHK(hipEventRecord(ev.e[0], s));
        for (int k = 0; k < tsweeps; ++k) {
          batch_sweep(s, a, c[0], c[1]);
          batch_next_sweep(s, ctr.p);
        }
        HIPCHK(hipEventRecord(ev.e[1], s));
        const double ms = ev.ms(h);
        HIPCHK(hipGetLastError());
        if (ms < best) {
          best = ms;
          block = c[0];
          grid = c[1];
        }
      }
      set_shape(grid);
      restore();
      HIPCHK(hipStreamSynchronize(s));
      tune_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else {
      set_shape(grid);
      restart();
    }

    // one sweep captured once (the kernel reads the sweep index from the device counter
    // that the second node bumps), replayed max_iter times
    if (log) {
      HIPCHK(hipMemsetAsync(lg.p, 0, lg.n * sizeof(BatchLog), s));
      a.log = lg.p;
    }
    HIPCHK(hipStreamSynchronize(s));
    GraphExec gx;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    batch_sweep(s, a, block, grid);
    batch_next_sweep(s, ctr.p);
    HIPCHK(hipStreamEndCapture(s, &gx.g));
    HIPCHK(hipGetLastError());
    HIPCHK(hipGraphInstantiate(&gx.e, gx.g, nullptr, nullptr, 0));
    TimingEvents ev(h);
    HIPCHK(hipEventRecord(ev.e[0], s));
    for (int it = 0; it < o.max_iter; ++it) HIPCHK(hipGraphLaunch(gx.e, s));
    HIPCHK(hipEventRecord(ev.e[1], s));
    const double ms = ev.ms(h);
    HIPCHK(hipGetLastError());

    std::vector<BatchProb> stf(P);
    std::vector<BatchDesc> df(P);
    int gfail = 0;
    HIPCHK(hipMemcpyAsync(&gfail, gerr.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(stf.data(), st.p, P * sizeof(BatchProb), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(df.data(), desc.p, P * sizeof(BatchDesc), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(rows.data(), fac.p, nr * sizeof(double), hipMemcpyDeviceToHost, s));
    if (log) HIPCHK(hipMemcpyAsync(log, lg.p, lg.n * sizeof(BatchLog), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    REQUIRE(!gfail, PTN_ERR_HIP, "a batch group's exchange wait ran out (members not co-resident?)");
    int64_t dst = 0;
    for (int p = 0; p < P; ++p) {
      const BatchDesc &d = b.desc[p];
      for (int k = 0; k < 3; ++k) {
        const int I = d.I[k];
        for (int r = 0; r < d.R; ++r)
          for (int i = 0; i < I; ++i) out_packed[dst + i + (int64_t)I * r] = rows[(d.f_off + d.rowoff[k] + i) * kBatchBR + r];
        dst += (int64_t)I * d.R;
      }
      ptn_result &rr = res[p];
      std::memset(&rr, 0, sizeof(rr));
      rr.frob_sq = df[p].frob;
      rr.f_value = stf[p].f;
      rr.rel_cost = stf[p].f / std::sqrt(df[p].frob);
      rr.ao_iter = stf[p].sweeps;
      rr.accel_coeff = stf[p].accel_coeff;
      rr.accel_fail = stf[p].accel_fail;
      rr.sweep_seconds = ms * 1e-3;
    }
    if (rep) {
      rep->block = block;
      rep->grid = grid;
      rep->tuned = tune ? 1 : 0;
      rep->seconds = ms * 1e-3;
      rep->tune_seconds = tune_s;
      rep->sweeps = o.max_iter;
      rep->split_problems = 0;
      rep->max_pieces = 1;
      for (const BatchDesc &d : dv) {
        rep->split_problems += d.V > 1;
        rep->max_pieces = std::max(rep->max_pieces, d.V);
      }
    }
  });
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// the warm-start network
// ------------------------------------------------------------------------------------
namespace {

// n host rows (FP32) into the activation buffers as bf16, zero rows up to the padded count
void upload_rows(MlpState &ms, int n, int np, const float *X, const float *T, const unsigned char *rank) {
  ptn_handle h = ms.h;
  std::vector<uint16_t> xb((size_t)np * kMlpIn, 0);
  for (size_t e = 0; e < (size_t)n * kMlpIn; ++e) xb[e] = f32_to_bf16(X[e]);
  HIPCHK(hipMemcpy(ms.X.p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
  if (T) {
    std::vector<float> t((size_t)np * kMlpR, 0.f);
    std::memcpy(t.data(), T, (size_t)n * kMlpR * sizeof(float));
    HIPCHK(hipMemcpy(ms.T.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<unsigned char> r((size_t)np, 0);
    std::memcpy(r.data(), rank, (size_t)n);
    HIPCHK(hipMemcpy(ms.rank.p, r.data(), r.size(), hipMemcpyHostToDevice));
  }
}

}  // namespace

extern "C" {

ptn_status ptn_mlp_init(ptn_handle h, uint64_t seed) {
  return run(h, [&] {
    MlpState &ms = mlp_of(h);
    std::vector<float> w(kMlpParams, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int64_t off, int64_t count, int fan_in) {
      std::uniform_real_distribution<double> u(-std::sqrt(6.0 / fan_in), std::sqrt(6.0 / fan_in));
      for (int64_t e = 0; e < count; ++e) w[off + e] = (float)u(rng);
    };
    fill(kMlpW1, (int64_t)kMlpH * kMlpIn, kMlpIn);
    fill(kMlpW2, (int64_t)kMlpH * kMlpH, kMlpH);
    fill(kMlpW3, (int64_t)kMlpR * kMlpH, kMlpH);
    HIPCHK(hipMemcpy(ms.w.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ms.m1.p, 0, kMlpParams * sizeof(float)));
    HIPCHK(hipMemset(ms.m2.p, 0, kMlpParams * sizeof(float)));
    weights_bf16(h->stream, kMlpParams, ms.w.p, ms.wb.p);
    HIPCHK(hipStreamSynchronize(h->stream));
    ms.step = 0;
    ms.ready = true;
  });
}

ptn_status ptn_mlp_set_params(ptn_handle h, const float *params) {
  return run(h, [&] {
    REQUIRE(params, PTN_ERR_INVALID, "params is null");
    MlpState &ms = mlp_of(h);
    HIPCHK(hipMemcpy(ms.w.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ms.m1.p, 0, kMlpParams * sizeof(float)));
    HIPCHK(hipMemset(ms.m2.p, 0, kMlpParams * sizeof(float)));
    weights_bf16(h->stream, kMlpParams, ms.w.p, ms.wb.p);
    HIPCHK(hipStreamSynchronize(h->stream));
    ms.step = 0;
    ms.ready = true;
  });
}

ptn_status ptn_mlp_get_params(ptn_handle h, float *params) {
  return run(h, [&] {
    REQUIRE(params, PTN_ERR_INVALID, "params is null");
    REQUIRE(h->mlp && h->mlp->ready, PTN_ERR_STATE, "no network on the handle");
    HIPCHK(hipMemcpy(params, h->mlp->w.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
  });
}

ptn_status ptn_mlp_forward(ptn_handle h, int n, const float *X, float *Y) {
  return run(h, [&] {
    REQUIRE(X && Y, PTN_ERR_INVALID, "null argument");
    REQUIRE(n >= 1 && n <= (1 << 22), PTN_ERR_INVALID, "n must be in [1, 2^22]");
    REQUIRE(h->mlp && h->mlp->ready, PTN_ERR_STATE, "no network on the handle");
    MlpState &ms = *h->mlp;
    ms.ensure_rows(n);
    upload_rows(ms, n, n, X, nullptr, nullptr);
    ms.forward(h->stream, n);
    HIPCHK(hipMemcpyAsync(Y, ms.Y.p, (size_t)n * kMlpR * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
  });
}

ptn_status ptn_mlp_loss_grad(ptn_handle h, int n, const float *X, const float *T, const unsigned char *rank,
                             double *loss, float *grad) {
  return run(h, [&] {
    REQUIRE(X && T && rank && loss && grad, PTN_ERR_INVALID, "null argument");
    REQUIRE(n >= 1 && n <= (1 << 20), PTN_ERR_INVALID, "n must be in [1, 2^20]");
    REQUIRE(h->mlp && h->mlp->ready, PTN_ERR_STATE, "no network on the handle");
    int64_t count = 0;
    for (int i = 0; i < n; ++i) {
      REQUIRE(rank[i] <= kMlpR, PTN_ERR_INVALID, "rank above 64");
      count += rank[i];
    }
    REQUIRE(count > 0, PTN_ERR_INVALID, "no valid entries");
    MlpState &ms = *h->mlp;
    const int np = (n + 31) / 32 * 32;
    ms.ensure_rows(np);
    upload_rows(ms, n, np, X, T, rank);
    ms.forward(h->stream, np);
    ms.backward(h->stream, np, (float)(1.0 / (double)count));
    std::vector<double> part(64);
    HIPCHK(hipMemcpyAsync(part.data(), ms.part.p, 64 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(grad, ms.g.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipGetLastError());
    double sum = 0.0;
    for (double v : part) sum += v;
    *loss = sum;
  });
}

ptn_status ptn_mlp_data_clear(ptn_handle h) {
  return run(h, [&] {
    MlpState &ms = mlp_of(h);
    ms.n = 0;
    ms.hrank.clear();
  });
}

ptn_status ptn_mlp_data_add(ptn_handle h, int n, const float *X, const float *T, const unsigned char *rank) {
  return run(h, [&] {
    REQUIRE(X && T && rank, PTN_ERR_INVALID, "null argument");
    REQUIRE(n >= 1 && n <= (1 << 24), PTN_ERR_INVALID, "n must be in [1, 2^24]");
    for (int i = 0; i < n; ++i) REQUIRE(rank[i] <= kMlpR, PTN_ERR_INVALID, "rank above 64");
    MlpState &ms = mlp_of(h);
    ms.ensure_data(ms.n + n);
    std::vector<uint16_t> xb((size_t)n * kMlpIn);
    for (size_t e = 0; e < xb.size(); ++e) xb[e] = f32_to_bf16(X[e]);
    HIPCHK(hipMemcpy(ms.dx.p + ms.n * kMlpIn, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ms.dt.p + ms.n * kMlpR, T, (size_t)n * kMlpR * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ms.drank.p + ms.n, rank, (size_t)n, hipMemcpyHostToDevice));
    ms.hrank.insert(ms.hrank.end(), rank, rank + n);
    ms.n += n;
  });
}

ptn_status ptn_mlp_data_size(ptn_handle h, int64_t *n) {
  return run(h, [&] {
    REQUIRE(n, PTN_ERR_INVALID, "null argument");
    *n = h->mlp ? h->mlp->n : 0;
  });
}

ptn_status ptn_mlp_data_get(ptn_handle h, int64_t first, int n, float *X, float *T, unsigned char *rank) {
  return run(h, [&] {
    REQUIRE(X && T && rank, PTN_ERR_INVALID, "null argument");
    REQUIRE(h->mlp && first >= 0 && n >= 0 && first + n <= h->mlp->n, PTN_ERR_INVALID, "rows out of range");
    MlpState &ms = *h->mlp;
    std::vector<uint16_t> xb((size_t)n * kMlpIn);
    HIPCHK(hipMemcpy(xb.data(), ms.dx.p + first * kMlpIn, xb.size() * 2, hipMemcpyDeviceToHost));
    for (size_t e = 0; e < xb.size(); ++e) {
      const uint32_t u = (uint32_t)xb[e] << 16;
      std::memcpy(&X[e], &u, 4);
    }
    HIPCHK(hipMemcpy(T, ms.dt.p + first * kMlpR, (size_t)n * kMlpR * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rank, ms.drank.p + first, (size_t)n, hipMemcpyDeviceToHost));
  });
}

ptn_status ptn_mlp_train(ptn_handle h, int epochs, int batch, double lr, uint64_t seed, double *epoch_loss) {
  return run(h, [&] {
    REQUIRE(h->mlp && h->mlp->ready, PTN_ERR_STATE, "no network on the handle");
    MlpState &ms = *h->mlp;
    REQUIRE(epochs >= 1 && epochs <= 100000, PTN_ERR_INVALID, "epochs must be in [1, 1e5]");
    REQUIRE(batch >= 32 && batch % 32 == 0 && batch <= 65536, PTN_ERR_INVALID,
            "batch must be a multiple of 32 in [32, 65536]");
    REQUIRE(ms.n >= batch, PTN_ERR_STATE, "the training set holds fewer rows than one batch");
    REQUIRE(lr > 0.0, PTN_ERR_INVALID, "lr must be positive");
    hipStream_t s = h->stream;
    ms.ensure_rows(batch);
    const int64_t spe = ms.n / batch;
    DevBuf<int> idx(h, (size_t)ms.n);
    DevBuf<double> losses(h, (size_t)(spe * epochs));
    std::vector<int> perm((size_t)ms.n);
    std::mt19937_64 rng(seed);
    for (int ep = 0; ep < epochs; ++ep) {
      std::iota(perm.begin(), perm.end(), 0);
      std::shuffle(perm.begin(), perm.end(), rng);
      HIPCHK(hipMemcpyAsync(idx.p, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, s));
      for (int64_t st = 0; st < spe; ++st) {
        int64_t count = 0;
        for (int64_t r = st * batch; r < (st + 1) * batch; ++r) count += ms.hrank[perm[r]];
        gather_rows(s, batch, idx.p + st * batch, ms.dx.p, ms.dt.p, ms.drank.p, ms.X.p, ms.T.p, ms.rank.p);
        ms.forward(s, batch);
        ms.backward(s, batch, count > 0 ? (float)(1.0 / (double)count) : 0.f);
        adam_step(s, kMlpParams, ms.w.p, ms.g.p, ms.m1.p, ms.m2.p, (float)lr, 0.9f, 0.999f, 1e-8f, ++ms.step, ms.wb.p);
        sum_partials(s, ms.part.p, 64, losses.p + ep * spe + st);
      }
      // the host permutation buffer is rewritten next epoch: wait for the upload to land
      HIPCHK(hipStreamSynchronize(s));
    }
    std::vector<double> l((size_t)(spe * epochs));
    HIPCHK(hipMemcpy(l.data(), losses.p, l.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipGetLastError());
    if (epoch_loss)
      for (int ep = 0; ep < epochs; ++ep) {
        double sm = 0.0;
        for (int64_t st = 0; st < spe; ++st) sm += l[ep * spe + st];
        epoch_loss[ep] = sm / (double)spe;
      }
  });
}

}  // extern "C"
