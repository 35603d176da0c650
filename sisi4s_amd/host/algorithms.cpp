// The hot-path algorithms with the reference's names, arguments, defaults
// and outputs, evaluated by the engine through the C ABI.
#include <cmath>
#include <sstream>

#include "algorithm.hpp"

namespace sisi_host {

namespace {

std::string num(double v) {
  std::ostringstream o;
  o.precision(17);
  o << v;
  return o.str();
}

std::shared_ptr<Tensor> make_tensor(std::vector<int64_t> lens) {
  auto t = std::make_shared<Tensor>();
  t->lens = std::move(lens);
  t->values.assign(size_t(t->size()), 0.0);
  return t;
}

sisi_engine *engine_of(Algorithm &a) {
  return engine_for(a.getTensorPtr("CoulombVertex"), a.getTensorPtr("HoleEigenEnergies"),
                    a.getTensorPtr("ParticleEigenEnergies"));
}

int precision_argument(Algorithm &a) {
  const std::string p = a.getTextArgument("precision", "fp64");
  if (p == "fp64" || p == "0") return SISI_FP64;
  if (p == "fp32" || p == "1") return SISI_FP32;
  if (p == "bf16x3" || p == "2") return SISI_BF16X3;
  throw StepError("unknown precision " + p + " (fp64, fp32, bf16x3)");
}

}  // namespace

// Seeded synthetic system (the engine's Philox generator); stands where the
// reference reads CoulombVertex / eigenenergies from files or generators.
class GenerateSyntheticSystem : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "GenerateSyntheticSystem"; }
  void run() override {
    const int No = int(getIntegerArgument("No")), Nv = int(getIntegerArgument("Nv"));
    const int NG = int(getIntegerArgument("NG"));
    const uint64_t seed = uint64_t(getIntegerArgument("seed", 20261202));
    auto g = make_tensor({NG, No + Nv, No + Nv});
    auto ei = make_tensor({No}), ea = make_tensor({Nv});
    check(sisi_generate_system(seed, No, Nv, NG, g->values.data(), ei->values.data(),
                               ea->values.data()),
          "sisi_generate_system");
    allocatedTensorArgument("CoulombVertex", g);
    allocatedTensorArgument("HoleEigenEnergies", ei);
    allocatedTensorArgument("ParticleEigenEnergies", ea);
    emitted.scalars.push_back({"No", std::to_string(No)});
    emitted.scalars.push_back({"Nv", std::to_string(Nv)});
    emitted.scalars.push_back({"NG", std::to_string(NG)});
  }
};

// UegVertexGenerator (UegVertexGenerator.cxx:49-229): same arguments and
// outputs; the vertex is written in real orbitals so that the real path
// applies (correlation energies equal the plane-wave ones).  NF must be left
// at its default (the reference's other values only serve profiling, cxx:162-163).
class UegVertexGenerator : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "UegVertexGenerator"; }
  void run() override {
    const int No = int(getIntegerArgument("No")), Nv = int(getIntegerArgument("Nv"));
    const double rs = getRealArgument("rs");
    const int hf = int(getIntegerArgument("hartreeFock", 1));
    const double madelung = getRealArgument("madelung", -1.0);
    const int64_t nf = getIntegerArgument("NF", 0);
    int NG = 0;
    check(sisi_generate_ueg(No, Nv, rs, hf, madelung, &NG, nullptr, nullptr, nullptr),
          "sisi_generate_ueg");
    if (nf != 0 && nf != NG)
      throw StepError("UegVertexGenerator: NF other than the transfer count (" +
                      std::to_string(NG) + ") is not supported");
    auto g = make_tensor({NG, No + Nv, No + Nv});
    auto ei = make_tensor({No}), ea = make_tensor({Nv});
    check(sisi_generate_ueg(No, Nv, rs, hf, madelung, &NG, g->values.data(), ei->values.data(),
                            ea->values.data()),
          "sisi_generate_ueg");
    allocatedTensorArgument("CoulombVertex", g);
    allocatedTensorArgument("HoleEigenEnergies", ei);
    allocatedTensorArgument("ParticleEigenEnergies", ea);
    emitted.scalars.push_back({"NF", std::to_string(NG)});
  }
};

// CoulombIntegralsFromVertex (CoulombIntegralsFromVertex.cxx:23-148, real
// integrals 203-571): publishes the requested classes.
class CoulombIntegralsFromVertex : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "CoulombIntegralsFromVertex"; }
  void run() override {
    if (getIntegerArgument("complex", 0) != 0) throw StepError("complex integrals are not supported");
    if (getIntegerArgument("antisymmetrize", 0) != 0)
      throw StepError("antisymmetrized integrals are not supported");
    sisi_engine *e = engine_of(*this);
    const auto &ei = getTensorArgument("HoleEigenEnergies");
    const auto &ea = getTensorArgument("ParticleEigenEnergies");
    const int64_t o = ei.lens[0], v = ea.lens[0];
    const struct {
      const char *name;
      int cls;
      std::vector<int64_t> lens;
    } classes[] = {{"PPHHCoulombIntegrals", SISI_PPHH, {v, v, o, o}},
                   {"PHPHCoulombIntegrals", SISI_PHPH, {v, o, v, o}},
                   {"HHHHCoulombIntegrals", SISI_HHHH, {o, o, o, o}},
                   {"HHHPCoulombIntegrals", SISI_HHHP, {o, o, o, v}},
                   {"HHPPCoulombIntegrals", SISI_HHPP, {o, o, v, v}}};
    for (auto &c : classes) {
      if (!out_.count(c.name)) continue;
      auto t = make_tensor(c.lens);
      check(sisi_get_integrals(e, c.cls, t->values.data()), "sisi_get_integrals");
      allocatedTensorArgument(c.name, t);
    }
    for (auto &kv : out_) {
      bool known = false;
      for (auto &c : classes) known |= (kv.first == c.name);
      if (!known) throw StepError("integral class " + kv.first + " is not built by this engine");
    }
    emitted.scalars.push_back({"No", std::to_string(o)});
    emitted.scalars.push_back({"Nv", std::to_string(v)});
  }
};

// CcsdEnergyFromCoulombIntegrals through ClusterSinglesDoublesAlgorithm::run
// (ClusterSinglesDoublesAlgorithm.cxx:25-128); argument names and defaults of
// ClusterSinglesDoublesAlgorithm.hpp:42-46, DiisMixer.cxx:55,
// LinearMixer.cxx:17, CcsdEnergyFromCoulombIntegrals.cxx:374-384.
class CcsdEnergyFromCoulombIntegrals : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "CcsdEnergyFromCoulombIntegrals"; }
  void run() override {
    if (isArgumentGiven("CoulombFactors")) throw StepError("CoulombFactors is not supported");
    if (getIntegerArgument("unrestricted", 0) != 0 || getIntegerArgument("antisymmetrize", 0) != 0)
      throw StepError("unrestricted / antisymmetrized energies are not supported");
    if (isArgumentGiven("HPFockMatrix")) throw StepError("HPFockMatrix is not supported");
    sisi_engine *e = engine_of(*this);
    const int64_t o = getTensorArgument("HoleEigenEnergies").lens[0];
    const int64_t v = getTensorArgument("ParticleEigenEnergies").lens[0];
    sisi_ccsd_opts opts;
    sisi_ccsd_opts_default(&opts);
    const std::string mixer = getTextArgument("mixer", "LinearMixer");
    if (mixer == "DiisMixer") opts.mixer = SISI_MIXER_DIIS;
    else if (mixer == "LinearMixer") opts.mixer = SISI_MIXER_LINEAR;
    else throw StepError("Mixer not implemented: " + mixer);
    opts.max_residua = int(getRealArgument("maxResidua", 4));
    opts.mixing_ratio = getRealArgument("mixingRatio", 1.0);
    opts.max_iterations = int(getIntegerArgument("maxIterations", 16));
    opts.amplitudes_convergence = getRealArgument("amplitudesConvergence", 1e-5);
    opts.energy_convergence = getRealArgument("energyConvergence", 1e-6);
    opts.level_shift = getRealArgument("levelShift", 0.0);
    int slice = int(getIntegerArgument("integralsSliceSize", -1));
    if (slice == -1 && isArgumentGiven("integralsSliceFactor"))
      slice = int(double(v) * getRealArgument("integralsSliceFactor"));
    opts.slice_size = slice;
    opts.precision = precision_argument(*this);
    opts.switch_threshold = getRealArgument("switchThreshold", 1e-5);
    opts.refine_interval = int(getIntegerArgument("refineInterval", 0));
    opts.ppl = getIntegerArgument("ppl", 1) != 0;                          // cxx:324
    opts.distinguishable = getIntegerArgument("distinguishable", 0) != 0;  // cxx:123-124
    {
      const std::string rp = getTextArgument("refinePrecision", "bf16");
      if (rp == "bf16" || rp == "3") opts.refine_precision = SISI_BF16;
      else if (rp == "bf16x3" || rp == "2") opts.refine_precision = SISI_BF16X3;
      else if (rp == "fp32" || rp == "1") opts.refine_precision = SISI_FP32;
      else throw StepError("unknown refinePrecision " + rp + " (bf16, bf16x3, fp32)");
    }
    std::vector<double> t1, t2;
    if (isArgumentGiven("initialSinglesAmplitudes")) t1 = getTensorArgument("initialSinglesAmplitudes").values;
    if (isArgumentGiven("initialDoublesAmplitudes")) t2 = getTensorArgument("initialDoublesAmplitudes").values;
    if (!t1.empty() && int64_t(t1.size()) != v * o) throw StepError("initialSinglesAmplitudes has the wrong size");
    if (!t2.empty() && int64_t(t2.size()) != v * v * o * o)
      throw StepError("initialDoublesAmplitudes has the wrong size");
    if (!t2.empty() && t1.empty()) t1.assign(size_t(v * o), 0.0);
    sisi_ccsd_report r;
    check(sisi_ccsd_solve(e, &opts, t1.empty() ? nullptr : t1.data(),
                          t2.empty() ? nullptr : t2.data(), &r),
          "sisi_ccsd_solve");
    int n = 0;
    check(sisi_ccsd_get_energies(e, nullptr, 0, &n), "sisi_ccsd_get_energies");
    std::vector<double> es(size_t(std::max(n, 1)));
    check(sisi_ccsd_get_energies(e, es.data(), n, &n), "sisi_ccsd_get_energies");
    for (int k = 0; k < n; ++k)
      emitted.iterations.push_back({{"iteration", std::to_string(k + 1)}, {"energy", num(es[k])}});
    emitted.scalars.push_back({"maxIterations", std::to_string(opts.max_iterations)});
    if (isArgumentGiven("distinguishable"))  // ClusterSinglesDoublesAlgorithm.cxx:69-71
      emitted.scalars.push_back({"distinguishable", std::to_string(getIntegerArgument("distinguishable"))});
    emitted.scalars.push_back({"converged", r.converged ? "true" : "false"});
    emitted.scalars.push_back({"fp64Iterations", std::to_string(r.fp64_iterations)});
    emitted.scalars.push_back({"energy", num(r.energy)});
    // energy decomposition of getEnergy (ClusterSinglesDoublesAlgorithm.cxx:180-190)
    emitted.scalars.push_back({"direct", num(r.direct)});
    emitted.scalars.push_back({"exchange", num(r.exchange)});
    emitted.scalars.push_back({"singlet", num(0.25 * r.direct - 0.5 * r.exchange)});
    emitted.scalars.push_back({"triplet", num(0.75 * r.direct + 1.5 * r.exchange)});
    if (isArgumentGiven("PairEnergy")) {  // cxx:218-235
      auto P = make_tensor({o, o});
      check(sisi_ccsd_pair_energies(e, P->values.data()), "sisi_ccsd_pair_energies");
      allocatedTensorArgument("PairEnergy", P);
    }
    if (isArgumentGiven("CcsdSinglesAmplitudes") || isArgumentGiven("CcsdDoublesAmplitudes")) {
      auto T1 = make_tensor({v, o}), T2 = make_tensor({v, v, o, o});
      check(sisi_ccsd_get_amplitudes(e, T1->values.data(), T2->values.data()),
            "sisi_ccsd_get_amplitudes");
      allocatedTensorArgument("CcsdSinglesAmplitudes", T1);
      allocatedTensorArgument("CcsdDoublesAmplitudes", T2);
    }
    setRealArgument("CcsdEnergy", r.energy);
  }
};

// CcsdPerturbativeTriples (CcsdPerturbativeTriples.cxx:119-248): output is
// CCSD energy + (T).
class CcsdPerturbativeTriples : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "CcsdPerturbativeTriples"; }
  void run() override {
    sisi_engine *e = engine_of(*this);
    const auto &T1 = getTensorArgument("CcsdSinglesAmplitudes");
    const auto &T2 = getTensorArgument("CcsdDoublesAmplitudes");
    const int64_t o = getTensorArgument("HoleEigenEnergies").lens[0];
    const int64_t v = getTensorArgument("ParticleEigenEnergies").lens[0];
    if (T1.size() != v * o || T2.size() != v * v * o * o)
      throw StepError("amplitudes do not match the eigenenergies");
    double et = 0;
    check(sisi_triples_energy(e, T1.values.data(), T2.values.data(), precision_argument(*this), &et),
          "sisi_triples_energy");
    const double ec = getRealArgument("CcsdEnergy");
    emitted.scalars.push_back({"triples", num(et)});
    emitted.scalars.push_back({"ccsd", num(ec)});
    emitted.scalars.push_back({"energy", num(ec + et)});
    setRealArgument("CcsdPerturbativeTriplesEnergy", ec + et);
  }
};

static Registrar<GenerateSyntheticSystem> r0("GenerateSyntheticSystem");
static Registrar<CoulombIntegralsFromVertex> r1("CoulombIntegralsFromVertex");
static Registrar<UegVertexGenerator> r4("UegVertexGenerator");
static Registrar<CcsdEnergyFromCoulombIntegrals> r2("CcsdEnergyFromCoulombIntegrals");
static Registrar<CcsdPerturbativeTriples> r3("CcsdPerturbativeTriples");

}  // namespace sisi_host
