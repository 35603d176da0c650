// The GEMM entry point that allocates nothing (sisi_gemm_device_noalloc).
//
// Same kernels as sisi_gemm_host (FP64 / FP32 MFMA, deterministic split-K),
// on device operands the caller owns and on the caller's stream.  The
// split-K partial sums go to `scratch`, a file-scope device array of
// kScratchDoubles doubles that every GPU the library runs on holds until
// the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress, on the entry point's first call with that device
// current) and only that address reaches kernels; the array's name in host
// code is a host-side stand-in, never a device address.
//
// Calls that overlap on one GPU share the array and race; the partial sums
// of one call are written in full before they are read, so nothing needs
// to start at zero.
#include <map>
#include <mutex>
#include <string>

#include "common.hpp"
#include "gemm.hpp"

namespace sisi {

void set_last_error(const std::string &msg);

namespace {

constexpr int64_t kScratchDoubles = int64_t(4) << 20;  // 32 MiB per device

__device__ __align__(256) double scratch[kScratchDoubles];

struct DeviceScratch {
  double *ptr = nullptr;
  int cus = 256;
};

std::mutex g_mu;
std::map<int, DeviceScratch> g_scratch;  // device ordinal -> address of its array

DeviceScratch scratch_of_current_device() {
  int dev = 0;
  SISI_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  DeviceScratch d;
  void *p = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  SISI_REQUIRE(rc == hipSuccess && p != nullptr, SISI_ERR_HIP,
               std::string("scratch address of device ") + std::to_string(dev) + ": " +
                   hipGetErrorString(rc));
  d.ptr = static_cast<double *>(p);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    d.cus = cus;
  g_scratch[dev] = d;
  return d;
}

}  // namespace
}  // namespace sisi

using namespace sisi;

extern "C" int sisi_gemm_device_noalloc(int precision, int transA, int transB, int M, int N,
                                        int K, double alpha, const double *A, int lda,
                                        const double *B, int ldb, double beta, double *C,
                                        int ldc, int ksplit, void *stream) {
  try {
    SISI_REQUIRE(precision == SISI_FP64 || precision == SISI_FP32, SISI_ERR_INVALID,
                 "gemm_device_noalloc: precision must be SISI_FP64 or SISI_FP32");
    SISI_REQUIRE(M >= 0 && N >= 0 && K >= 0 && ksplit >= 0, SISI_ERR_INVALID,
                 "gemm_device_noalloc: negative dimension or split");
    const int64_t ra = transA ? K : M, rb = transB ? N : K;
    SISI_REQUIRE(lda >= std::max<int64_t>(ra, 1) && ldb >= std::max<int64_t>(rb, 1) &&
                     ldc >= std::max(M, 1),
                 SISI_ERR_INVALID, "gemm_device_noalloc: leading dimension too small");
    if (M > 0 && N > 0)
      SISI_REQUIRE(C != nullptr && (K == 0 || (A != nullptr && B != nullptr)), SISI_ERR_INVALID,
                   "gemm_device_noalloc: null pointer");
    const DeviceScratch d = scratch_of_current_device();
    GemmArgs g;
    g.A = A;
    g.B = B;
    g.C = C;
    g.M = M;
    g.N = N;
    g.K = K;
    g.sam = transA ? lda : 1;
    g.sak = transA ? 1 : lda;
    g.sbk = transB ? ldb : 1;
    g.sbn = transB ? 1 : ldb;
    g.cm0 = 1;
    g.cn0 = ldc;
    g.dm0 = std::max(M, 1);
    g.dn0 = std::max(N, 1);
    g.alpha = alpha;
    g.beta = beta;
    gemm_workspace(static_cast<hipStream_t>(stream), d.cus, precision, g, d.ptr, kScratchDoubles,
                   ksplit);
    set_last_error("");
    return SISI_OK;
  } catch (const Error &err) {
    set_last_error(err.what());
    return err.code;
  } catch (const std::exception &ex) {
    set_last_error(ex.what());
    return SISI_ERR_INVALID;
  } catch (...) {
    set_last_error("unknown error");
    return SISI_ERR_INVALID;
  }
}
