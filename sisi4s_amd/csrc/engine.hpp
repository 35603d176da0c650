// Engine state: device copies of the inputs, integrals, CCSD iteration state
// and the kernel-family timers.  One engine = one HIP stream on the device
// that was current at creation.
#pragma once

#include <array>
#include <map>
#include <string>
#include <memory>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"

namespace sisi {

enum class Kind { Gemm64 = 0, Gemm32 = 1, GemmB3 = 2, Bandwidth = 3 };

// Records (start, stop) event pairs around launches while stats are on.
class Timers {
 public:
  explicit Timers(hipStream_t s) : stream_(s) {}
  ~Timers();
  void begin(Kind k, double work);
  void end();
  void collect(sisi_stats &out);
  void reset();
  bool enabled = false;

 private:
  struct Rec {
    Kind kind;
    double work;
    hipEvent_t a, b;
  };
  hipEvent_t take();
  hipStream_t stream_;
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> spare_;
  sisi_stats acc_{};
  bool open_ = false;
};

// Scoped timing of one launch (no-op when stats are off).
class Timed {
 public:
  Timed(Timers &t, Kind k, double work) : t_(t) {
    if (t_.enabled) t_.begin(k, work);
  }
  ~Timed() {
    if (t_.enabled) t_.end();
  }

 private:
  Timers &t_;
};

// FP64 evaluation point of the refinement iterations: amplitudes, residual
// and the T-dependent intermediates of that residual.
struct Anchor {
  bool valid = false;
  int age = 0;  // refinement iterations since the anchor was set
  int precision = SISI_FP64;  // of the anchor residual (FP64, or a bulk precision)
  bool ppl = true, distinguishable = false;  // residual variant it was recorded with
  DBuf T1, T2, R1, R2;
  std::map<std::string, DBuf> v;
  // bf16 planes of the ladder's V~(T_A) slice pairs (pair order m <= n),
  // split from the FP64 slices the anchor computes; an empty store marks a
  // pair too small for the bf16 kernels
  std::vector<B3FrozenA> ladder;
  int ladder_slice = 0;
  // when every pair is large enough, their planes are stacked in one store
  // (pair p at rows ladder_off[p], ladder_rows rows in all, lo plane after
  // the hi plane): the refinement contracts all pairs with dX in one product
  DBuf ladder_stack;
  int64_t ladder_rows = 0;
  std::vector<int64_t> ladder_off;
  // bf16 hi/lo planes of S[c,k,b,d] = sum_G Gamma_ck^G dressed Gamma_bd^G(T_A)
  // as rows k + No b, columns c + Nv d: the change of the ladder vertex is
  // linear in the T1 change through Gamma_ck, so with S the refinement forms
  // dV~ X(T) from one product instead of one per slice pair
  B3FrozenA ladder_s;
};

struct Tape;

struct DiisSlot {
  DBuf t1, t2;   // estimated amplitudes
  DBuf d1, d2;   // amplitude change
};

class Engine {
 public:
  explicit Engine(const sisi_dims &d);
  ~Engine();

  int No, Nv, NG, Np;
  int device = 0;
  int cus = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  Pool pool;
  std::unique_ptr<Timers> timers;

  // Deferred launches (lockstep batched solver): while set, the launch
  // helpers append records to the tape instead of launching
  Tape *tape = nullptr;

  // precision of general contractions and of the particle-particle ladder
  int precision = SISI_FP64;
  int prec_ladder = SISI_FP64;
  // residual variant (CcsdEnergyFromCoulombIntegrals.cxx:123-124, 324-338):
  // ppl = false drops the particle-particle ladder, distinguishable selects
  // the distinguishable-cluster terms
  bool ppl = true, distinguishable = false;
  // Reduced-precision ladder (split-BF16x3 / bf16): the V~_xycd product
  // writes the bf16 planes the X contraction reads (no FP64 V~ slice pair
  // and no operand split of it); off: the FP64 slice pair as in the FP64
  // ladder, split by the next product
  bool ladder_planes = true;

  // Operands marked constant for the duration of one residual evaluation:
  // the contraction planner caches their permuted dense copies.
  std::vector<const double *> frozen;
  std::map<std::string, DBuf> perm_cache;
  std::map<std::string, DBuf> split_cache;  // bf16 hi/lo planes of frozen operands
  bool is_frozen(const double *p) const {
    for (auto q : frozen)
      if (q == p) return true;
    return false;
  }
  void thaw() {
    frozen.clear();
    perm_cache.clear();
    split_cache.clear();
  }
  // Forget one frozen operand and every cached copy of it (before its buffer
  // goes back to the pool, whose next user must not inherit the caches).
  void unfreeze(const double *p) {
    for (size_t k = 0; k < frozen.size();)
      if (frozen[k] == p) frozen.erase(frozen.begin() + k);
      else ++k;
    const std::string tag = std::to_string(reinterpret_cast<uintptr_t>(p));
    for (auto *cache : {&perm_cache, &split_cache})
      for (auto it = cache->begin(); it != cache->end();) {
        const std::string &k = it->first;  // "<ptr>:..." or "b3:<ptr>:..."
        const bool hit = k.compare(0, tag.size() + 1, tag + ":") == 0 ||
                         k.compare(0, tag.size() + 4, "b3:" + tag + ":") == 0;
        if (hit) it = cache->erase(it);
        else ++it;
      }
  }

  // Complex mode (vertex given by sisi_engine_set_vertex_complex): the
  // vertex, its slices, the integrals, amplitudes and residuals are stored
  // planar, every real part followed by the imaginary part of the same
  // tensor, so sizes double and host arrays are converted from/to
  // interleaved complex128 at the boundary.
  bool cplx = false;

  // inputs
  DBuf gamma, eps_i, eps_a;
  DBuf Gij, Gai, Gab;  // dense slices [NG,No,No], [NG,Nv,No], [NG,Nv,Nv]
  bool have_vertex = false, have_eps = false, have_integrals = false;

  // integrals
  DBuf Vabij, Vaibj, Vijkl, Vijka;
  // complex mode only: Gamma_ia slice and the classes the complex residual
  // reads besides the ones above (CoulombIntegralsFromVertex.cxx:762-810);
  // U_klcd = 2 V_klcd - V_kldc of HHPP; HHPP as [a,b,i,j] and with a,b
  // exchanged (energy terms)
  DBuf Gia, Vaijb, Vijab, Vaijk, Uklcd, VijabT, VjiabT;
  // Coulomb factors (CoulombFactors Lambda[G,R], FactorOrbitals Pi[p,R],
  // planar complex) and V_RS = sum_G conj(Lambda_GR) Lambda_GS; NR > 0 makes
  // the ladder the factor-based one (CcsdEnergyFromCoulombIntegrals.cxx:340-372)
  // Full-integral formulation (CcsdEnergyFromCoulombIntegralsReference):
  // PPPP V_abcd and PPPH V_abci, and which classes were given directly
  // (sisi_engine_set_integrals, no vertex needed); formulation 0: vertex
  // residual, 1: full-integral residual, 2: its onlyPPL branch
  DBuf Vabcd, Vabci;
  unsigned given_integrals = 0;
  int formulation = 0;
  // Direct-ring CCD (formulation SISI_FORMULATION_DRCCD): spins = 1 for
  // `unrestricted`, the linearized and adjacentPairsExchange arguments, and
  // uploaded PHHP / HHPP integrals (derived from PPHH when absent)
  double spins = 2.0;
  // getEnergy variants (ClusterSinglesDoublesAlgorithm.cxx:131-213): the
  // antisymmetrized energy (antisymmetrize: 1) and the HPFockMatrix term
  // spins * sum_ai T_ai f_ia, f stored transposed as fai[a,i] (planar
  // complex in complex mode); `noncanonical` is its last value
  bool anti_energy = false;
  DBuf fai;
  bool fai_cplx = false;
  double noncanonical = 0.0;
  bool ring_linearized = false, ring_adjacent = false;
  DBuf ring_Vaijb, ring_Vijab;
  int NR = 0, factors_slice = 0;
  bool factors_real = true;  // imaginary parts all zero: real products only
  DBuf Lam, PiF, PiO, VRS;  // PiO: OutgoingFactorOrbitals (complex amplitudes)
  DBuf Vbaij;  // V_abij with a,b exchanged (energy exchange term)
  DBuf Uabij;  // 2 V_abij - V_baij
  // (T) integrals, built on first use: V_bdck as [d,b,c,k] and as [d,c,b,k],
  // V_jklc as [l,c,j,k]
  DBuf Vt, Vt2, VP;
  // the same operands as split-BF16x3 planes (V_jklc with the sign of its
  // term folded in), built on the first reduced-precision (T) call
  B3Planes Vt_b3, Vt2_b3, VP_b3;
  // EOM-CCSD: the amplitudes the similarity-transformed Hamiltonian is built
  // from and its blocks, built once per amplitude set (IP and EA blocks on
  // their first right-apply) and dropped with the integrals
  struct EomHbar {
    bool have_t = false, ip = false, ea = false;
    std::vector<double> hT1, hT2;  // host copies of the amplitudes set
    DBuf T1, T2, Fme, Foo, Fvv, Woooo, Wovvo, Wooov, Wovoo, Wvvvv, Wvovv, Wvvvo;
  } eom;
  void drop_eom() { eom = EomHbar(); }

  void drop_triples_cache() {
    drop_eom();
    Vt.reset();
    Vt2.reset();
    VP.reset();
    Vt_b3 = B3Planes();
    Vt2_b3 = B3Planes();
    VP_b3 = B3Planes();
  }

  // refinement anchor (recorded by an FP64 residual when `recording` is set)
  Anchor anchor;
  bool recording = false;
  void record(const std::string &name, const double *p, size_t n);

  // CCSD run state
  sisi_ccsd_opts opts{};
  bool run_active = false;
  bool run_initial = false;
  int iteration = 0;
  int fp64_iterations = 0;
  int cur_precision = SISI_FP64;
  bool coarse_active = false;  // mixed run still in its coarse stage
  double prev_energy = 0.0, last_energy = 0.0, last_direct = 0.0,
         last_exchange = 0.0, last_delta = 0.0;
  double prev_energy_imag = 0.0, last_energy_imag = 0.0;  // complex mode
  bool converged = false;
  // refinement phase: an anchor asked for by a refinement iteration that met
  // the convergence test is pending; after one that did not confirm it, the
  // next confirm_hold refinement iterations do not ask again
  bool confirm_pending = false;
  int confirm_hold = 0;
  DBuf T1, T2;
  std::vector<double> energies;
  // DIIS ring (DiisMixer.cxx:48-180)
  std::vector<DiisSlot> diis;
  std::vector<std::vector<double>> B;  // (N+1)x(N+1)
  int diis_next = 0, diis_count = 0;
  // linear mixer (LinearMixer.cxx:30-55) keeps only the last vector
  DBuf lin1, lin2;
  bool lin_have = false;

  // helpers
  void require_inputs() const;
  void upload(DBuf &dst, const double *src, size_t n);
  void download(double *dst, const DBuf &src, size_t n);
  // n doubles of a tensor as the caller holds it: in complex mode the host
  // array is interleaved complex128 and the device copy planar
  void upload_t(DBuf &dst, const double *src, size_t n);
  void download_t(double *dst, const DBuf &src, size_t n);
  void sync();
  // doubles per singles / doubles amplitude tensor (both planes in complex
  // mode) and elements per plane
  size_t n1() const { return n1e() * (cplx ? 2 : 1); }
  size_t n2() const { return n2e() * (cplx ? 2 : 1); }
  size_t n1e() const { return size_t(Nv) * No; }
  size_t n2e() const { return size_t(Nv) * Nv * No * No; }

  // hipBLASLt state of the split-BF16x3 library backend (handle, workspace,
  // algorithms per shape); declared last so it is released before the pool
  std::shared_ptr<void> blaslt;

  // Residual captured in a hipGraph (opts.graphs): fixed input/output
  // buffers, the pool buffers the graph references, and what it was captured
  // for (precision, slice size, formulation, variants, factors)
  struct ResidualGraph {
    hipGraphExec_t exec = nullptr;
    DBuf T1, T2, R1, R2;
    std::vector<void *> held;
    std::vector<int> key;
  } rgraph;
  void drop_residual_graph();
};

}  // namespace sisi
