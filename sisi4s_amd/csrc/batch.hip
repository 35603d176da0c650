// Lockstep batched CCSD: many independent systems iterated together, every
// launch of an iteration shared by all systems.
//
// The reference runs one system per driver invocation
// (ClusterSinglesDoublesAlgorithm.cxx:38-128).  For a batch of small systems
// a per-system residual is a few hundred launches of a few microseconds
// each, so the device idles between launches whatever the number of host
// threads.  Here each system's residual is recorded once per precision as a
// tape of launches (engine.tape), the tapes of all systems are merged into
// one schedule of grouped launches (GroupedPlan: one launch per step and
// kernel variant, each workgroup finding its system from a prefix sum of
// block counts), and the schedule is replayed every iteration, captured in a
// hipGraph once it has been reused.  The update of each iteration (amplitude
// estimate, DIIS overlaps, the DIIS combination, the convergence norms and
// the energy) is recorded per system and merged the same way; the host reads
// the reductions of all systems back at once, solves the small DIIS systems
// in FP64 and checks convergence per system.  Launches per iteration
// therefore do not grow with the number of systems.
//
// Per system the iteration follows ClusterSinglesDoublesAlgorithm.cxx:85-128
// with the mixers of DiisMixer.cxx:103-180 and LinearMixer.cxx:30-44,
// exactly as the single-system driver does; the reduced precisions of a
// mixed run (bulk and coarse stages) run their products on FP32 MFMA, the
// remaining iterations in FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "ccsd.hpp"
#include "engine.hpp"
#include "tape.hpp"
#include "tensor_ops.hpp"

namespace sisi {

namespace {

// Byte size of the argument record of a tape op.
size_t record_bytes(OpKind k) {
  if (k == OpKind::Gemm || k == OpKind::GemmReduce) return sizeof(GemmArgs);
  if (k == OpKind::CopyTiled || k == OpKind::CopyGeneric) return sizeof(CopyArgs);
  return sizeof(ElemArgs);
}

const void *record_of(const TapeOp &op) {
  if (op.kind == OpKind::Gemm || op.kind == OpKind::GemmReduce) return &op.g;
  if (op.kind == OpKind::CopyTiled || op.kind == OpKind::CopyGeneric) return &op.c;
  return &op.e;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Device buffer of `bytes` bytes from the pool with host bytes copied in
// (stream-ordered on s, host-synchronous).
void upload_bytes(Pool &pool, hipStream_t s, const std::vector<unsigned char> &host, DBuf &out) {
  out = DBuf(&pool, std::max<size_t>(1, (host.size() + 7) / 8));
  if (!host.empty())
    SISI_HIP(hipMemcpyAsync(out.get(), host.data(), host.size(), hipMemcpyHostToDevice, s));
  SISI_HIP(hipStreamSynchronize(s));
}

int key_of(const TapeOp &op) { return int(op.kind) * 16 + op.sub; }

}  // namespace

void Tape::upload(Pool &pool, hipStream_t s) {
  std::vector<unsigned char> host;
  off.assign(ops.size(), 0);
  for (size_t i = 0; i < ops.size(); ++i) {
    const size_t b = record_bytes(ops[i].kind);
    off[i] = align_up(host.size(), 16);
    host.resize(off[i] + b);
    std::memcpy(host.data() + off[i], record_of(ops[i]), b);
  }
  upload_bytes(pool, s, host, dev);
}

void GroupedPlan::reset() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  exec_ = nullptr;
  groups_.clear();
  dev_.reset();
  runs_ = 0;
  records_ = 0;
}

void GroupedPlan::build(Pool &pool, const std::vector<const Tape *> &tapes, hipStream_t s) {
  reset();
  // Greedy merge: at each step launch the (kind, variant) whose next records
  // carry the most remaining work (sum of the remaining tape lengths of the
  // tapes whose next record has that key), which keeps the long tapes moving.
  constexpr int kKeys = 16 * 16;
  const size_t nt = tapes.size();
  std::vector<size_t> pos(nt, 0);
  struct Member {
    int tape;
    size_t op;
  };
  std::vector<std::pair<int, std::vector<Member>>> merged;
  for (;;) {
    double score[kKeys] = {};
    bool any = false;
    for (size_t t = 0; t < nt; ++t) {
      const auto &ops = tapes[t]->ops;
      if (pos[t] >= ops.size()) continue;
      any = true;
      score[key_of(ops[pos[t]])] += double(ops.size() - pos[t]);
    }
    if (!any) break;
    const int best = int(std::max_element(score, score + kKeys) - score);
    std::vector<Member> m;
    for (size_t t = 0; t < nt; ++t) {
      const auto &ops = tapes[t]->ops;
      if (pos[t] < ops.size() && key_of(ops[pos[t]]) == best) m.push_back({int(t), pos[t]++});
    }
    merged.emplace_back(best, std::move(m));
  }
  // Layout: argument records of the tapes without resident records, then
  // per group its pointer array and its block starts.
  std::vector<unsigned char> host;
  std::vector<std::vector<size_t>> packed(nt);
  for (size_t t = 0; t < nt; ++t) {
    const Tape &tp = *tapes[t];
    if (tp.dev.get()) continue;
    packed[t].resize(tp.ops.size());
    for (size_t i = 0; i < tp.ops.size(); ++i) {
      const size_t b = record_bytes(tp.ops[i].kind);
      packed[t][i] = align_up(host.size(), 16);
      host.resize(packed[t][i] + b);
      std::memcpy(host.data() + packed[t][i], record_of(tp.ops[i]), b);
    }
  }
  for (auto &mg : merged) {
    Group g;
    g.kind = OpKind(mg.first / 16);
    g.sub = mg.first % 16;
    g.n = int(mg.second.size());
    g.ptrs = align_up(host.size(), 16);
    g.start = g.ptrs + sizeof(void *) * size_t(g.n);
    g.gxs = g.start + sizeof(unsigned) * size_t(g.n + 1);
    host.resize(g.gxs + (g.kind == OpKind::GemmReduce ? sizeof(unsigned) * size_t(g.n) : 0));
    groups_.push_back(g);
  }
  records_ = 0;
  for (auto &mg : merged) records_ += mg.second.size();
  dev_ = DBuf(&pool, std::max<size_t>(1, (host.size() + 7) / 8));
  const uintptr_t base = reinterpret_cast<uintptr_t>(dev_.get());
  for (size_t gi = 0; gi < merged.size(); ++gi) {
    Group &g = groups_[gi];
    uint64_t total = 0;
    for (int p = 0; p < g.n; ++p) {
      const Member &m = merged[gi].second[size_t(p)];
      const Tape &tp = *tapes[size_t(m.tape)];
      const TapeOp &op = tp.ops[m.op];
      const uintptr_t rec = tp.dev.get()
                                ? reinterpret_cast<uintptr_t>(tp.dev.get()) + tp.off[m.op]
                                : base + packed[size_t(m.tape)][m.op];
      std::memcpy(host.data() + g.ptrs + sizeof(void *) * size_t(p), &rec, sizeof(rec));
      const unsigned st = unsigned(total);
      std::memcpy(host.data() + g.start + sizeof(unsigned) * size_t(p), &st, sizeof(st));
      if (g.kind == OpKind::GemmReduce)
        std::memcpy(host.data() + g.gxs + sizeof(unsigned) * size_t(p), &op.gx, sizeof(op.gx));
      total += uint64_t(op.blocks());
    }
    SISI_REQUIRE(total > 0 && total < (uint64_t(1) << 31), SISI_ERR_INVALID,
                 "grouped launch: block count out of range");
    if (profile_shapes) {
      const TapeOp *big = nullptr;
      for (const Member &m : merged[gi].second) {
        const TapeOp &op = tapes[size_t(m.tape)]->ops[m.op];
        if (!big || op.blocks() > big->blocks()) big = &op;
      }
      char buf[160];
      if (g.kind == OpKind::Gemm || g.kind == OpKind::GemmReduce)
        snprintf(buf, sizeof(buf), "%s/%d M=%d N=%d K=%d b=%d ks=%d", g.kind == OpKind::Gemm ? "gemm" : "reduce",
                 g.sub, big->g.M, big->g.N, big->g.K, big->g.batch, big->g.ksplit);
      else if (g.kind == OpKind::CopyTiled || g.kind == OpKind::CopyGeneric)
        snprintf(buf, sizeof(buf), "copy/%d n=%lld nd=%d", int(g.kind), (long long)big->c.total,
                 big->c.nd);
      else
        snprintf(buf, sizeof(buf), "elem/%d n=%lld", int(g.kind), (long long)big->e.n);
      g.label = buf;
    }
    const unsigned st = unsigned(total);
    std::memcpy(host.data() + g.start + sizeof(unsigned) * size_t(g.n), &st, sizeof(st));
    g.total = st;
  }
  if (!host.empty())
    SISI_HIP(hipMemcpyAsync(dev_.get(), host.data(), host.size(), hipMemcpyHostToDevice, s));
  SISI_HIP(hipStreamSynchronize(s));
}

void GroupedPlan::run(hipStream_t s, bool graph) {
  auto launch_one = [&](const Group &g) {
    const unsigned char *base = reinterpret_cast<const unsigned char *>(dev_.get());
    {
      const unsigned *start = reinterpret_cast<const unsigned *>(base + g.start);
      if (g.kind == OpKind::Gemm)
        launch_gemm_grouped(g.sub, reinterpret_cast<const GemmArgs *const *>(base + g.ptrs), start,
                            g.n, g.total, s);
      else if (g.kind == OpKind::GemmReduce)
        launch_gemm_reduce_grouped(reinterpret_cast<const GemmArgs *const *>(base + g.ptrs),
                                   start, reinterpret_cast<const unsigned *>(base + g.gxs), g.n,
                                   g.total, s);
      else if (g.kind == OpKind::CopyTiled || g.kind == OpKind::CopyGeneric)
        launch_copy_grouped(g.kind, reinterpret_cast<const CopyArgs *const *>(base + g.ptrs),
                            start, g.n, g.total, s);
      else
        launch_elem_grouped(g.kind, reinterpret_cast<const ElemArgs *const *>(base + g.ptrs),
                            start, g.n, g.total, s);
    }
  };
  auto launch_all = [&]() {
    for (const Group &g : groups_) launch_one(g);
  };
  if (profile) {
    std::vector<hipEvent_t> ev(2 * groups_.size());
    for (auto &x : ev) SISI_HIP(hipEventCreate(&x));
    for (size_t i = 0; i < groups_.size(); ++i) {
      SISI_HIP(hipEventRecord(ev[2 * i], s));
      launch_one(groups_[i]);
      SISI_HIP(hipEventRecord(ev[2 * i + 1], s));
    }
    SISI_HIP(hipStreamSynchronize(s));
    profile->resize(256);
    for (size_t i = 0; i < groups_.size(); ++i) {
      float ms = 0.f;
      SISI_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      ProfileEntry &p = (*profile)[size_t(int(groups_[i].kind) * 16 + groups_[i].sub)];
      p.ms += double(ms);
      p.launches += 1.0;
      p.records += double(groups_[i].n);
      p.blocks += double(groups_[i].total);
      if (profile_shapes) {
        ProfileEntry &q = (*profile_shapes)[groups_[i].label];
        q.ms += double(ms);
        q.launches += 1.0;
        q.records += double(groups_[i].n);
        q.blocks += double(groups_[i].total);
      }
    }
    for (auto &x : ev) (void)hipEventDestroy(x);
    ++runs_;
    return;
  }
  if (exec_) {
    SISI_HIP(hipGraphLaunch(exec_, s));
  } else if (graph && runs_ >= 2) {
    hipGraph_t gr = nullptr;
    hipError_t err = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
    if (err == hipSuccess) {
      try {
        launch_all();
      } catch (...) {
        (void)hipStreamEndCapture(s, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        (void)hipGetLastError();
        throw;
      }
      err = hipStreamEndCapture(s, &gr);
    }
    if (err == hipSuccess) err = hipGraphInstantiate(&exec_, gr, nullptr, nullptr, 0);
    if (gr) (void)hipGraphDestroy(gr);
    if (err == hipSuccess) {
      SISI_HIP(hipGraphLaunch(exec_, s));
    } else {
      (void)hipGetLastError();
      exec_ = nullptr;
      launch_all();
    }
  } else {
    launch_all();
  }
  ++runs_;
}

namespace {

// One system of the batch: engine, fixed amplitude and residual buffers
// (the recorded residuals read and write them), mixer state and run state.
struct System {
  std::unique_ptr<Engine> e;
  DBuf T1, T2, R1, R2, d1, d2;  // d: amplitude change (linear mixer)
  DBuf part, out;               // reduction partials and results
  std::vector<DiisSlot> slots;
  std::vector<std::vector<double>> B;
  int next = 0, count = 0;
  bool lin_have = false;
  Tape res[2];                   // residual: [0] FP64, [1] reduced precision (FP32)
  std::vector<void *> held[2];   // pool buffers the recorded residual uses
  Tape step;                     // per-iteration update records
  // run state (ClusterSinglesDoublesAlgorithm.cxx:85-128)
  int iteration = 0, fp64_iterations = 0;
  int cur_precision = SISI_FP64;
  bool coarse_active = false, converged = false, active = true;
  double prev_energy = 0.0, energy = 0.0, direct = 0.0, exchange = 0.0, delta = 1.0;
  int status = SISI_OK;
  std::string error;
};

constexpr int kOutSlots = 64;  // doubles of reduction results per system

// Residual of system s at precision index p recorded on its tape.
void record_residual(System &s, int p, int slice, int it, Tape &tape, std::vector<void *> *held,
                     Pool &pool, hipStream_t stream) {
  Engine &e = *s.e;
  set_precision(e, p == 0 ? SISI_FP64 : SISI_FP32);
  tape.clear();
  e.tape = &tape;
  e.pool.begin_capture();
  try {
    ccsd_residual(e, it, false, slice, s.T1.get(), s.T2.get(), s.R1.get(), s.R2.get());
  } catch (...) {
    e.tape = nullptr;
    e.pool.release(e.pool.end_capture());
    throw;
  }
  e.tape = nullptr;
  std::vector<void *> h = e.pool.end_capture();
  if (held) *held = std::move(h);
  else e.pool.release(h);
  tape.upload(pool, stream);
}

}  // namespace

void ccsd_solve_lockstep(int nsys, const sisi_dims *dims, const double *const *gammas,
                         const double *const *eps_i, const double *const *eps_a,
                         const sisi_ccsd_opts &o, double *energies, sisi_ccsd_report *reports,
                         sisi_lockstep_stats *stats) {
  SISI_REQUIRE(o.mixer == SISI_MIXER_LINEAR || o.mixer == SISI_MIXER_DIIS, SISI_ERR_INVALID,
               "Mixer not implemented");
  SISI_REQUIRE(o.mixer != SISI_MIXER_DIIS || (o.max_residua >= 1 && o.max_residua <= 30),
               SISI_ERR_INVALID, "maxResidua must be in [1, 30] for the lockstep solver");
  SISI_REQUIRE(o.refine_interval == 0, SISI_ERR_INVALID,
               "the lockstep solver runs plain FP64 iterations after the switch "
               "(refine_interval 0)");
  SISI_REQUIRE(o.bulk_refine_interval == 0, SISI_ERR_INVALID,
               "the lockstep solver runs plain bulk iterations (bulk_refine_interval 0)");
  SISI_REQUIRE(is_precision(o.precision), SISI_ERR_INVALID, "unknown precision");
  SISI_REQUIRE(o.coarse_threshold <= 0.0 || is_precision(o.coarse_precision), SISI_ERR_INVALID,
               "coarse_precision must be a precision");
  using clock = std::chrono::steady_clock;
  const auto t_start = clock::now();
  auto since = [](clock::time_point t0) {
    return std::chrono::duration<double>(clock::now() - t0).count();
  };
  Pool pool;
  hipStream_t stream = nullptr;
  SISI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  // device time of the grouped residual / update launches (events read after
  // the read-back that ends every update)
  hipEvent_t ev[4] = {};
  struct StreamGuard {
    hipStream_t s;
    hipEvent_t *ev;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      for (int k = 0; k < 4; ++k)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
      (void)hipStreamDestroy(s);
    }
  } guard{stream, ev};
  for (int k = 0; k < 4; ++k) SISI_HIP(hipEventCreate(&ev[k]));
  double res_ms = 0.0, upd_ms = 0.0;
  auto elapsed = [&](int a, int b) {
    float ms = 0.f;
    SISI_HIP(hipEventElapsedTime(&ms, ev[a], ev[b]));
    return double(ms);
  };
  std::vector<System> sys(static_cast<size_t>(nsys));
  const int N = o.mixer == SISI_MIXER_DIIS ? o.max_residua : 0;
  const bool bulk_run = o.precision != SISI_FP64;
  // set-up: integrals per system (eager, on the system's own stream), fixed
  // buffers, recorded residuals
  for (int k = 0; k < nsys; ++k) {
    System &s = sys[size_t(k)];
    s.e = std::make_unique<Engine>(dims[k]);
    Engine &e = *s.e;
    const size_t np = size_t(dims[k].No) + dims[k].Nv;
    e.upload(e.gamma, gammas[k], size_t(dims[k].NG) * np * np);
    e.have_vertex = true;
    e.upload(e.eps_i, eps_i[k], size_t(dims[k].No));
    e.upload(e.eps_a, eps_a[k], size_t(dims[k].Nv));
    e.have_eps = true;
    integrals_from_vertex(e);
    e.ppl = o.ppl != 0;
    e.distinguishable = o.distinguishable != 0;
    const size_t n1 = e.n1(), n2 = e.n2();
    s.T1 = DBuf(&pool, n1);
    s.T2 = DBuf(&pool, n2);
    s.R1 = DBuf(&pool, n1);
    s.R2 = DBuf(&pool, n2);
    s.d1 = DBuf(&pool, n1);
    s.d2 = DBuf(&pool, n2);
    s.part = DBuf(&pool, 2 * kTapeRedBlocks);
    s.slots.resize(size_t(N));
    if (N > 0) {
      s.B.assign(size_t(N + 1), std::vector<double>(size_t(N + 1), 0.0));
      for (int i = 1; i <= N; ++i) s.B[0][size_t(i)] = s.B[size_t(i)][0] = -1.0;  // DiisMixer.cxx:83-96
    }
    s.cur_precision = o.precision;
    s.coarse_active = bulk_run && o.coarse_threshold > 0.0;
    if (bulk_run) record_residual(s, 1, o.slice_size, 1, s.res[1], &s.held[1], pool, stream);
    record_residual(s, 0, o.slice_size, 1, s.res[0], &s.held[0], pool, stream);
  }
  if (stats) stats->setup_seconds = since(t_start);
  // reduction results of all systems in one array (one read-back)
  DBuf outs(&pool, size_t(std::max(nsys, 1)) * kOutSlots);
  std::vector<double> hout(size_t(std::max(nsys, 1)) * kOutSlots);
  auto out_of = [&](int k) { return outs.get() + size_t(k) * kOutSlots; };
  auto read_outs = [&]() {
    SISI_HIP(hipMemcpyAsync(hout.data(), outs.get(), hout.size() * sizeof(double),
                            hipMemcpyDeviceToHost, stream));
    SISI_HIP(hipStreamSynchronize(stream));
  };
  auto run_tapes = [&](const std::vector<const Tape *> &tapes, GroupedPlan &plan) {
    if (tapes.empty()) return;
    plan.build(pool, tapes, stream);
    plan.run(stream, false);
    if (stats) {
      stats->launches += double(plan.launches());
      stats->records += double(plan.records());
    }
  };
  // residual schedule: rebuilt when a system changes precision, or when a
  // quarter of the systems it runs have stopped (stopped systems keep
  // running in the schedule until then; their results are not read)
  GroupedPlan res_plan, step_plan;
  // diagnostics: SISI_LOCKSTEP_PROFILE=1 times every grouped launch of the
  // residual (eagerly, no graph) and prints the time per kernel kind to stderr
  std::vector<GroupedPlan::ProfileEntry> prof;
  const char *penv = getenv("SISI_LOCKSTEP_PROFILE");
  std::map<std::string, GroupedPlan::ProfileEntry> prof_shapes;
  if (penv && atoi(penv) > 0) {
    res_plan.profile = &prof;
    res_plan.profile_shapes = &prof_shapes;
  }
  std::vector<std::pair<int, int>> res_members;  // (system, precision index)
  const double shift = o.level_shift;
  for (int iter = 0;; ++iter) {
    std::vector<int> act;
    for (int k = 0; k < nsys; ++k)
      if (sys[size_t(k)].active) act.push_back(k);
    if (act.empty()) break;
    auto prec_index = [&](const System &s) {
      const bool bulk = bulk_run && s.cur_precision != SISI_FP64;
      return bulk ? 1 : 0;
    };
    // residual
    if (iter == 0) {
      // MP2 start (CcsdEnergyFromCoulombIntegrals.cxx:47-51): R1 = 0, R2 = V
      std::vector<Tape> t0(act.size());
      std::vector<const Tape *> tp;
      for (size_t i = 0; i < act.size(); ++i) {
        System &s = sys[size_t(act[i])];
        s.e->tape = &t0[i];
        // T = 0 before the first estimate (which reads T: level shift, dT)
        zero(*s.e, s.T1.get(), int64_t(s.e->n1()));
        zero(*s.e, s.T2.get(), int64_t(s.e->n2()));
        ccsd_residual(*s.e, 0, false, o.slice_size, s.T1.get(), s.T2.get(), s.R1.get(),
                      s.R2.get());
        s.e->tape = nullptr;
        tp.push_back(&t0[i]);
      }
      run_tapes(tp, step_plan);
    } else {
      bool rebuild = res_members.empty();
      int stopped = 0;
      for (auto &m : res_members) {
        const System &s = sys[size_t(m.first)];
        if (!s.active) ++stopped;
        else if (prec_index(s) != m.second) rebuild = true;
      }
      if (4 * stopped > int(res_members.size())) rebuild = true;
      if (rebuild) {
        res_members.clear();
        std::vector<const Tape *> tp;
        for (int k : act) {
          const int p = prec_index(sys[size_t(k)]);
          res_members.push_back({k, p});
          tp.push_back(&sys[size_t(k)].res[p]);
        }
        res_plan.build(pool, tp, stream);
        if (stats) stats->plans += 1.0;
      }
      SISI_HIP(hipEventRecord(ev[0], stream));
      res_plan.run(stream, true);
      SISI_HIP(hipEventRecord(ev[1], stream));
      if (stats) {
        stats->launches += double(res_plan.launches());
        stats->records += double(res_plan.records());
      }
    }
    // update, part 1: estimate and the mixer's overlaps
    SISI_HIP(hipEventRecord(ev[2], stream));
    std::vector<const Tape *> tp;
    for (int k : act) {
      System &s = sys[size_t(k)];
      Engine &e = *s.e;
      const int64_t n1 = int64_t(e.n1()), n2 = int64_t(e.n2());
      s.step.clear();
      e.tape = &s.step;
      double *out = out_of(k);
      if (N > 0) {
        DiisSlot &sl = s.slots[size_t(s.next)];
        if (!sl.t2.get()) {
          sl.t1 = DBuf(&pool, size_t(n1));
          sl.t2 = DBuf(&pool, size_t(n2));
          sl.d1 = DBuf(&pool, size_t(n1));
          sl.d2 = DBuf(&pool, size_t(n2));
        }
        copy(e, dense(sl.t1.get(), {n1}), dense(s.R1.get(), {n1}));
        copy(e, dense(sl.t2.get(), {n2}), dense(s.R2.get(), {n2}));
        estimate_amplitudes(e, 2, sl.t1.get(), s.T1.get(), sl.d1.get(), shift);
        estimate_amplitudes(e, 4, sl.t2.get(), s.T2.get(), sl.d2.get(), shift);
        for (int i = 0; i < N; ++i) {
          const DiisSlot &si = s.slots[size_t(i)];
          if (!si.d2.get()) continue;
          dot_into(e, n1, si.d1.get(), sl.d1.get(), s.part.get(), out + 2 * i);
          dot_into(e, n2, si.d2.get(), sl.d2.get(), s.part.get(), out + 2 * i + 1);
        }
      } else {
        // LinearMixer::append (LinearMixer.cxx:30-44): the previous mixed
        // amplitudes are the current T
        estimate_amplitudes(e, 2, s.R1.get(), s.T1.get(), s.d1.get(), shift);
        estimate_amplitudes(e, 4, s.R2.get(), s.T2.get(), s.d2.get(), shift);
        if (s.lin_have) {
          axpby(e, n1, 1.0 - o.mixing_ratio, s.T1.get(), o.mixing_ratio, s.R1.get());
          axpby(e, n2, 1.0 - o.mixing_ratio, s.T2.get(), o.mixing_ratio, s.R2.get());
        }
        s.lin_have = true;
        copy(e, dense(s.T1.get(), {n1}), dense(s.R1.get(), {n1}));
        copy(e, dense(s.T2.get(), {n2}), dense(s.R2.get(), {n2}));
        dot_into(e, n1, s.d1.get(), s.d1.get(), s.part.get(), out + 0);
        dot_into(e, n2, s.d2.get(), s.d2.get(), s.part.get(), out + 1);
        dot_into(e, n1, s.T1.get(), s.T1.get(), s.part.get(), out + 2);
        dot_into(e, n2, s.T2.get(), s.T2.get(), s.part.get(), out + 3);
        energy_into(e, s.T1.get(), s.T2.get(), s.part.get(), out + 4);
      }
      e.tape = nullptr;
      tp.push_back(&s.step);
    }
    run_tapes(tp, step_plan);
    SISI_HIP(hipEventRecord(ev[3], stream));
    read_outs();
    if (iter > 0) {
      const double ms = elapsed(0, 1);
      res_ms += ms;
      if (res_plan.profile && (iter % 10 == 0 || act.size() < 4))
        fprintf(stderr, "[lockstep iteration] %d: %zu systems, %zu launches, residual %.2f ms\n",
                iter, act.size(), res_plan.launches(), ms);
    }
    upd_ms += elapsed(2, 3);
    // update, part 2 (DIIS): combination, norms and energy
    if (N > 0) {
      tp.clear();
      for (int k : act) {
        System &s = sys[size_t(k)];
        Engine &e = *s.e;
        const int64_t n1 = int64_t(e.n1()), n2 = int64_t(e.n2());
        const double *h = hout.data() + size_t(k) * kOutSlots;
        for (int i = 0; i < N; ++i) {
          if (!s.slots[size_t(i)].d2.get()) continue;
          const double ov = 2.0 * (h[2 * i] + h[2 * i + 1]);
          s.B[size_t(s.next + 1)][size_t(i + 1)] = ov;
          s.B[size_t(i + 1)][size_t(s.next + 1)] = ov;
        }
        s.delta = h[2 * s.next] + h[2 * s.next + 1];  // |dT|^2, normalised below
        if (s.count < N) ++s.count;
        const int dim = s.count + 1;
        std::vector<std::vector<double>> Bm(static_cast<size_t>(dim), std::vector<double>(static_cast<size_t>(dim)));
        for (int r = 0; r < dim; ++r)
          for (int c = 0; c < dim; ++c) Bm[size_t(r)][size_t(c)] = s.B[size_t(r)][size_t(c)];
        std::vector<double> rhs(size_t(dim), 0.0);
        rhs[0] = -1.0;
        std::vector<double> col;
        try {
          col = diis_solve(Bm, rhs);
        } catch (const Error &err) {
          s.status = err.code;
          s.error = err.what();
          s.active = false;
          continue;
        }
        s.step.clear();
        e.tape = &s.step;
        zero(e, s.T1.get(), n1);
        zero(e, s.T2.get(), n2);
        for (int j = 0; j < s.count; ++j) {
          const int i = (s.next + N - j) % N;
          axpby(e, n1, col[size_t(i + 1)], s.slots[size_t(i)].t1.get(), 1.0, s.T1.get());
          axpby(e, n2, col[size_t(i + 1)], s.slots[size_t(i)].t2.get(), 1.0, s.T2.get());
        }
        s.next = (s.next + 1) % N;
        double *out = out_of(k);
        dot_into(e, n1, s.T1.get(), s.T1.get(), s.part.get(), out + 2 * N);
        dot_into(e, n2, s.T2.get(), s.T2.get(), s.part.get(), out + 2 * N + 1);
        energy_into(e, s.T1.get(), s.T2.get(), s.part.get(), out + 2 * N + 2);
        e.tape = nullptr;
        tp.push_back(&s.step);
      }
      SISI_HIP(hipEventRecord(ev[2], stream));
      run_tapes(tp, step_plan);
      SISI_HIP(hipEventRecord(ev[3], stream));
      read_outs();
      upd_ms += elapsed(2, 3);
    }
    // energies, convergence and the precision schedule per system
    for (int k : act) {
      System &s = sys[size_t(k)];
      if (!s.active) continue;
      const double *h = hout.data() + size_t(k) * kOutSlots;
      double dd, tt, rd, rx;
      if (N > 0) {
        dd = s.delta;
        tt = h[2 * N] + h[2 * N + 1];
        rd = h[2 * N + 2];
        rx = h[2 * N + 3];
      } else {
        dd = h[0] + h[1];
        tt = h[2] + h[3];
        rd = h[4];
        rx = h[5];
      }
      s.delta = std::sqrt(std::fabs(dd / tt));
      // getEnergy (ClusterSinglesDoublesAlgorithm.cxx:131-189), spins = 2
      s.direct = 2.0 * rd;
      s.exchange = -rx;
      s.energy = s.direct + s.exchange;
      const bool bulk = bulk_run && s.cur_precision != SISI_FP64;
      if (!std::isfinite(s.energy) || !std::isfinite(s.delta)) {
        s.status = SISI_ERR_NUMERIC;
        s.error = "CCSD iteration " + std::to_string(s.iteration + 1) +
                  " diverged (non-finite energy or amplitudes)";
        s.active = false;
        continue;
      }
      if (!bulk) s.fp64_iterations++;
      s.iteration++;
      if (s.coarse_active && s.delta <= o.coarse_threshold && s.iteration > 1)
        s.coarse_active = false;
      if (bulk && !s.coarse_active && s.delta <= o.switch_threshold && s.iteration > 1)
        s.cur_precision = SISI_FP64;
      const double de = std::fabs((s.energy - s.prev_energy) / s.energy);
      const bool met = de < std::fabs(o.energy_convergence) &&
                       s.delta * s.delta <
                           std::fabs(o.amplitudes_convergence * o.amplitudes_convergence);
      if (met && !bulk) s.converged = true;
      s.prev_energy = s.energy;
      if (s.converged || s.iteration >= o.max_iterations) s.active = false;
    }
    if (stats) stats->iterations += 1.0;
  }
  SISI_HIP(hipStreamSynchronize(stream));
  if (stats) {
    stats->residual_seconds = 1e-3 * res_ms;
    stats->update_seconds = 1e-3 * upd_ms;
  }
  for (int k = 0; k < nsys; ++k) {
    System &s = sys[size_t(k)];
    energies[k] = s.status == SISI_OK ? s.energy : std::nan("");
    if (reports) {
      sisi_ccsd_report &r = reports[k];
      r = sisi_ccsd_report{};
      r.energy = energies[k];
      r.direct = s.direct;
      r.exchange = s.exchange;
      r.delta = s.delta;
      r.iterations = s.iteration;
      r.converged = s.converged ? 1 : 0;
      r.fp64_iterations = s.fp64_iterations;
      // the reduced precisions of a lockstep run are FP32 products
      r.precision = (bulk_run && s.cur_precision != SISI_FP64) ? SISI_FP32 : SISI_FP64;
      r.status = s.status;
      r.energy_imag = 0.0;
    }
  }
  // release the recorded residuals' buffers before the engines go
  for (auto &s : sys)
    for (int p = 0; p < 2; ++p) {
      s.e->pool.release(s.held[p]);
      s.held[p].clear();
    }
  if (stats) stats->total_seconds = since(t_start);
  if (!prof.empty()) {
    static const char *kinds[] = {"gemm", "copy_tiled", "copy_generic", "fill", "axpby",
                                  "add_t1t1", "estimate", "dot_partial", "sum_partials",
                                  "energy_partial", "gemm_reduce"};
    fprintf(stderr, "[lockstep profile] kind/variant: ms, launches, records, blocks\n");
    for (size_t k = 0; k < prof.size(); ++k)
      if (prof[k].launches > 0)
        fprintf(stderr, "[lockstep profile] %s/%d: %.1f ms, %.0f, %.0f, %.0f\n",
                k / 16 < 11 ? kinds[k / 16] : "?", int(k % 16), prof[k].ms, prof[k].launches,
                prof[k].records, prof[k].blocks);
    std::vector<std::pair<double, std::string>> top;
    for (auto &kv : prof_shapes) top.push_back({kv.second.ms, kv.first});
    std::sort(top.rbegin(), top.rend());
    for (size_t i = 0; i < top.size() && i < 40; ++i) {
      const auto &q = prof_shapes[top[i].second];
      fprintf(stderr, "[lockstep shape] %8.1f ms %7.0f launches %8.0f records %10.0f blocks  %s\n",
              q.ms, q.launches, q.records, q.blocks, top[i].second.c_str());
    }
  }
  for (int k = 0; k < nsys; ++k)
    if (sys[size_t(k)].status != SISI_OK)
      throw Error(sys[size_t(k)].status,
                  "system " + std::to_string(k) + ": " + sys[size_t(k)].error);
}

}  // namespace sisi
