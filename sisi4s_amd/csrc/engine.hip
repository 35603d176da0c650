// Engine construction, host<->device transfers and launch timers.
#include "engine.hpp"

#include <cmath>

namespace sisi {

Timers::~Timers() {
  for (auto &r : recs_) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto ev : spare_) (void)hipEventDestroy(ev);
}

hipEvent_t Timers::take() {
  if (!spare_.empty()) {
    hipEvent_t ev = spare_.back();
    spare_.pop_back();
    return ev;
  }
  hipEvent_t ev;
  SISI_HIP(hipEventCreate(&ev));
  return ev;
}

void Timers::begin(Kind k, double work) {
  Rec r{k, work, take(), take()};
  SISI_HIP(hipEventRecord(r.a, stream_));
  recs_.push_back(r);
  open_ = true;
}

void Timers::end() {
  if (!open_ || recs_.empty()) return;
  SISI_HIP(hipEventRecord(recs_.back().b, stream_));
  open_ = false;
  // fold finished records now and then so long runs keep few events alive
  if (recs_.size() > 4096) {
    sisi_stats tmp{};
    collect(tmp);
  }
}

void Timers::collect(sisi_stats &out) {
  SISI_HIP(hipStreamSynchronize(stream_));
  for (auto &r : recs_) {
    float ms = 0.f;
    SISI_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    const double s = ms * 1e-3;
    switch (r.kind) {
      case Kind::Gemm64:
      case Kind::Gemm32:
      case Kind::GemmB3: {
        const int k = int(r.kind);
        acc_.gemm_seconds[k] += s;
        acc_.gemm_flops[k] += r.work;
        acc_.gemm_launches[k] += 1;
        break;
      }
      case Kind::Bandwidth:
        acc_.other_seconds += s;
        acc_.other_bytes += r.work;
        acc_.other_launches += 1;
        break;
    }
    spare_.push_back(r.a);
    spare_.push_back(r.b);
  }
  recs_.clear();
  out = acc_;
}

void Timers::reset() {
  sisi_stats tmp{};
  collect(tmp);
  acc_ = sisi_stats{};
}

Engine::Engine(const sisi_dims &d) {
  SISI_REQUIRE(d.No > 0 && d.Nv > 0 && d.NG > 0, SISI_ERR_INVALID,
               "dims: No, Nv, NG must be positive");
  SISI_REQUIRE(int64_t(d.Nv) * d.Nv * d.No * d.No < (int64_t(1) << 40), SISI_ERR_INVALID,
               "dims: amplitudes too large");
  No = d.No;
  Nv = d.Nv;
  NG = d.NG;
  Np = d.No + d.Nv;
  int count = 0;
  SISI_HIP(hipGetDeviceCount(&count));
  SISI_REQUIRE(count > 0, SISI_ERR_HIP, "no HIP device");
  SISI_HIP(hipGetDevice(&device));
  SISI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  if (cus <= 0) cus = 256;
  {
    std::lock_guard<std::recursive_mutex> lock(device_wide_mutex());
    SISI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  }
  timers.reset(new Timers(stream));
}

void Engine::drop_residual_graph() {
  if (rgraph.exec) {
    (void)hipStreamSynchronize(stream);
    (void)hipGraphExecDestroy(rgraph.exec);
    rgraph.exec = nullptr;
  }
  pool.release(rgraph.held);
  rgraph.held.clear();
  rgraph.key.clear();
  rgraph.T1.reset();
  rgraph.T2.reset();
  rgraph.R1.reset();
  rgraph.R2.reset();
}

Engine::~Engine() {
  std::lock_guard<std::recursive_mutex> lock(device_wide_mutex());
  (void)hipStreamSynchronize(stream);
  drop_residual_graph();
  timers.reset();
  // device buffers return to the pool before it is released
  T1.reset();
  T2.reset();
  anchor = Anchor();
  diis.clear();
  lin1.reset();
  lin2.reset();
  perm_cache.clear();
  split_cache.clear();
  gamma.reset();
  eps_i.reset();
  eps_a.reset();
  Gij.reset();
  Gai.reset();
  Gab.reset();
  Vabij.reset();
  Vaibj.reset();
  Vijkl.reset();
  Vijka.reset();
  Vbaij.reset();
  Uabij.reset();
  drop_triples_cache();
  blaslt.reset();  // library handle and workspace, under the device-wide lock
  pool.release_all();
  if (stream) (void)hipStreamDestroy(stream);
}

void Engine::require_inputs() const {
  SISI_REQUIRE(have_vertex, SISI_ERR_STATE, "no CoulombVertex set");
  SISI_REQUIRE(have_eps, SISI_ERR_STATE, "no eigenenergies set");
}

void Engine::upload(DBuf &dst, const double *src, size_t n) {
  SISI_REQUIRE(src != nullptr, SISI_ERR_INVALID, "null host pointer");
  if (dst.size() < n) dst = DBuf(&pool, n);
  SISI_HIP(hipMemcpyAsync(dst.get(), src, n * sizeof(double), hipMemcpyHostToDevice, stream));
  SISI_HIP(hipStreamSynchronize(stream));
}

void Engine::download(double *dst, const DBuf &src, size_t n) {
  SISI_REQUIRE(dst != nullptr, SISI_ERR_INVALID, "null host pointer");
  SISI_REQUIRE(src.get() != nullptr && src.size() >= n, SISI_ERR_STATE, "no such device data");
  SISI_HIP(hipMemcpyAsync(dst, src.get(), n * sizeof(double), hipMemcpyDeviceToHost, stream));
  SISI_HIP(hipStreamSynchronize(stream));
}

void Engine::upload_t(DBuf &dst, const double *src, size_t n) {
  if (!cplx) return upload(dst, src, n);
  SISI_REQUIRE(src != nullptr, SISI_ERR_INVALID, "null host pointer");
  const size_t h = n / 2;
  std::vector<double> planar(n);
  for (size_t k = 0; k < h; ++k) {
    planar[k] = src[2 * k];
    planar[h + k] = src[2 * k + 1];
  }
  upload(dst, planar.data(), n);
}

void Engine::download_t(double *dst, const DBuf &src, size_t n) {
  if (!cplx) return download(dst, src, n);
  SISI_REQUIRE(dst != nullptr, SISI_ERR_INVALID, "null host pointer");
  const size_t h = n / 2;
  std::vector<double> planar(n);
  download(planar.data(), src, n);
  for (size_t k = 0; k < h; ++k) {
    dst[2 * k] = planar[k];
    dst[2 * k + 1] = planar[h + k];
  }
}

void Engine::sync() { SISI_HIP(hipStreamSynchronize(stream)); }

void Engine::record(const std::string &name, const double *p, size_t n) {
  if (!recording) return;
  DBuf &b = anchor.v[name];
  if (b.size() != n) b = DBuf(&pool, n);
  SISI_HIP(hipMemcpyAsync(b.get(), p, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

}  // namespace sisi
