// Linear-solver-level entry points of the C ABI: Jacobi-preconditioned CG on
// a caller's symmetric positive definite 3x3-block CSR matrix, with the
// semantics of the reference's "eigen-pcg-diag" solver (Eigen's
// ConjugateGradient<SpMat, Lower|Upper> with its default DiagonalPreconditioner,
// reference src/linear_solvers/eigen_iterative_solver.h:14-29, configured by
// src/factories/linear_solver_factory.cpp:75-85). The iteration itself is the
// library's one-kernel-per-iteration CG (k_solver.hip), in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/mfem_hip.h"
#include "host/abi_error.h"
#include "host/device.h"
#include "kernels.h"

namespace mfemhip {
namespace {

constexpr int kCheckBS = 256;

__global__ void k_zero_flag(int* flag) { *flag = 0; }

// Pattern check and Eigen's DiagonalPreconditioner::factorize (1 / a_jj, or 1
// where the diagonal entry is zero or absent; the first diagonal block of a
// row counts). One block row per lane. flag bit 0: rowptr[0] != 0, a row
// pointer out of order or past nnzb; bit 1: a block column outside [0, nb).
__global__ __launch_bounds__(kCheckBS) void k_bsr_check_invdiag(
    int nb, int nnzb, const int* __restrict__ rowptr, const int* __restrict__ col,
    const double* __restrict__ K, double* __restrict__ invdiag, int* flag) {
  const int i = blockIdx.x * kCheckBS + threadIdx.x;
  if (i >= nb) return;
  const int p0 = rowptr[i], p1 = rowptr[i + 1];
  double d[3] = {1.0, 1.0, 1.0};
  int bad = (i == 0 && p0 != 0) ? 1 : 0;
  if (p0 < 0 || p1 < p0 || p1 > nnzb) {
    bad |= 1;
  } else {
    bool found = false;
    for (int b = p0; b < p1; ++b) {
      const int c = col[b];
      if (c < 0 || c >= nb) {
        bad |= 2;
      } else if (c == i && !found) {
        found = true;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const double v = K[9L * b + 4 * r];
          d[r] = v != 0.0 ? 1.0 / v : 1.0;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) invdiag[3L * i + r] = d[r];
  if (bad) atomicOr(flag, bad);
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// Byte layout of the work arrays of one solve on n = 3 nb scalar rows: the
// control block, the pattern flag, the two partial-sum buffers, D^-1 and the
// two CG state buffers ({r, q, p, D^-1} per row).
struct PcgWorkLayout {
  size_t ctrl = 0, flag = 0, parts = 0, invdiag = 0, state0 = 0, state1 = 0, total = 0;
  explicit PcgWorkLayout(long n) {
    size_t o = 0;
    ctrl = o;
    o += align256(sizeof(PcgCtrl));
    flag = o;
    o += align256(sizeof(int));
    parts = o;
    o += align256(sizeof(double) * 2 * kCgParts * (size_t)kCgMaxBlocks);
    invdiag = o;
    o += align256(sizeof(double) * (size_t)n);
    state0 = o;
    o += align256(sizeof(double) * 4 * (size_t)n);
    state1 = o;
    o += align256(sizeof(double) * 4 * (size_t)n);
    total = o;
  }
};

size_t pcg_bsr3_work_bytes(long n) { return PcgWorkLayout(n).total; }

// The solve on work arrays at `work` (PcgWorkLayout(3 nb).total bytes of
// device memory). Synchronous: returns after the result is on the host.
void pcg_bsr3_run(char* work, int nb, int nnzb, const int* rowptr, const int* col,
                  const double* vals, const double* b, double* x, double tol, int max_iters,
                  int* iterations, double* rel_residual, hipStream_t st) {
  if (nb < 0 || nnzb < 0) throw std::invalid_argument("negative matrix size");
  if (!(tol >= 0.0)) throw std::invalid_argument("tolerance must be >= 0");
  if (max_iters < 0) throw std::invalid_argument("max_iters must be >= 0");
  if (nb == 0) {
    if (iterations) *iterations = 0;
    if (rel_residual) *rel_residual = 0.0;
    return;
  }
  if (!rowptr || !col || !vals || !b || !x || !work)
    throw std::invalid_argument("null device pointer");
  if (3L * nb > (long)(1 << 30)) throw std::invalid_argument("matrix too large");
  const int n = 3 * nb;
  const PcgWorkLayout L(n);
  int* flag = reinterpret_cast<int*>(work + L.flag);
  double* invdiag = reinterpret_cast<double*>(work + L.invdiag);
  hipLaunchKernelGGL(k_zero_flag, dim3(1), dim3(1), 0, st, flag);
  hipLaunchKernelGGL(k_bsr_check_invdiag, dim3((nb + kCheckBS - 1) / kCheckBS), dim3(kCheckBS), 0,
                     st, nb, nnzb, rowptr, col, vals, invdiag, flag);
  MFEM_HIP_CHECK(hipGetLastError());
  int hflag = 0;
  MFEM_HIP_CHECK(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  if (hflag & 1) throw std::invalid_argument("malformed block row pointer (rowptr)");
  if (hflag & 2) throw std::invalid_argument("block column index out of range");

  PcgBuffers<double> B;
  B.n = n;
  B.rowptr = rowptr;
  B.col = col;
  B.K = vals;
  B.invdiag = invdiag;
  B.b = b;
  B.x = x;
  B.state[0] = reinterpret_cast<double*>(work + L.state0);
  B.state[1] = reinterpret_cast<double*>(work + L.state1);
  B.ctrl = reinterpret_cast<PcgCtrl*>(work + L.ctrl);
  B.parts[0] = reinterpret_cast<double*>(work + L.parts);
  B.parts[1] = B.parts[0] + (size_t)kCgParts * kCgMaxBlocks;
  B.variant = 0;
  B.blocks = cg_blocks_per_cu(0, n, 4);
  launch_pcg_init(B, tol, max_iters, st);
  // kernels 0 .. max_iters + 1 at most; the control block is read back every
  // 16 launches, and launches queued after the stop exit at once
  const long limit = (long)max_iters + 2;
  long launched = 0;
  PcgCtrl h{};
  while (true) {
    const long chunk = std::min<long>(16, limit - launched);
    for (long j = 0; j < chunk; ++j) launch_pcg_iteration(B, (int)(launched + j), st);
    launched += chunk;
    MFEM_HIP_CHECK(hipGetLastError());
    MFEM_HIP_CHECK(hipMemcpyAsync(&h, B.ctrl, sizeof(PcgCtrl), hipMemcpyDeviceToHost, st));
    MFEM_HIP_CHECK(hipStreamSynchronize(st));
    if (h.done) break;
    if (launched >= limit) throw std::runtime_error("CG control block did not terminate");
  }
  if (iterations) *iterations = h.iter;
  if (rel_residual) *rel_residual = h.b2 > 0 ? std::sqrt(h.rr / h.b2) : 0.0;
}

}  // namespace mfemhip

extern "C" int mfem_pcg_bsr3(int nb, int nnzb, const int* rowptr, const int* col,
                             const double* vals, const double* b, double* x, double tol,
                             int max_iters, int* iterations, double* rel_residual, void* stream) {
  return mfem::abi_guarded([&] {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* work = nullptr;
    if (nb > 0) {
      const size_t bytes = mfemhip::pcg_bsr3_work_bytes(3L * nb);
      MFEM_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&work), bytes));
    }
    try {
      mfemhip::pcg_bsr3_run(work, nb, nnzb, rowptr, col, vals, b, x, tol, max_iters, iterations,
                            rel_residual, st);
    } catch (...) {
      if (work) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(work);
      }
      throw;
    }
    if (work) MFEM_HIP_CHECK(hipFree(work));
  });
}

// ---------------------------------------------------------------------------
// The same solve without any device allocation: the work arrays live in a
// fixed array of this library's code object, one per GPU, sized for
// MFEM_PCG_NOALLOC_MAX_ROWS scalar rows. In host code `scratch` is only a
// stand-in symbol; its device address comes from hipGetSymbolAddress, once per
// device, and only that address is handed to kernels.
namespace mfemhip {
namespace {

constexpr long kNoAllocRows = MFEM_PCG_NOALLOC_MAX_ROWS;
constexpr size_t kNoAllocBytes =
    sizeof(double) * 9 * (size_t)kNoAllocRows + sizeof(double) * 2 * kCgParts * kCgMaxBlocks +
    4 * 256;
static_assert(kNoAllocBytes % sizeof(double) == 0, "scratch size");

static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call there). Called with g_noalloc_mu held.
char* scratch_on_current_device() {
  int dev = 0;
  MFEM_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                               hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace mfemhip

extern "C" int mfem_pcg_bsr3_noalloc(int nb, int nnzb, const int* rowptr, const int* col,
                                     const double* vals, const double* b, double* x, double tol,
                                     int max_iters, int* iterations, double* rel_residual,
                                     void* stream) {
  return mfem::abi_guarded([&] {
    if (nb < 0) throw std::invalid_argument("negative matrix size");
    if (3L * nb > mfemhip::kNoAllocRows)
      throw std::invalid_argument(
          "system of " + std::to_string(3L * nb) + " rows exceeds the no-allocation limit of " +
          std::to_string(mfemhip::kNoAllocRows) + " rows (MFEM_PCG_NOALLOC_MAX_ROWS)");
    if (mfemhip::pcg_bsr3_work_bytes(3L * nb) > mfemhip::kNoAllocBytes)
      throw std::logic_error("no-allocation work layout exceeds the scratch array");
    std::lock_guard<std::mutex> lock(mfemhip::g_noalloc_mu);
    char* work = nb > 0 ? mfemhip::scratch_on_current_device() : nullptr;
    mfemhip::pcg_bsr3_run(work, nb, nnzb, rowptr, col, vals, b, x, tol, max_iters, iterations,
                          rel_residual, reinterpret_cast<hipStream_t>(stream));
  });
}
