// The hot-path algorithms with the reference's names, arguments, defaults
// and outputs, evaluated by the engine through the C ABI.
#include <cmath>
#include <fstream>
#include <sstream>

#include "algorithm.hpp"
#include "davidson.hpp"
#include "eom_space.hpp"

namespace sisi_host {

namespace {

std::string num(double v) {
  std::ostringstream o;
  o.precision(17);
  o << v;
  return o.str();
}

std::shared_ptr<Tensor> make_tensor(std::vector<int64_t> lens, bool complex = false) {
  auto t = std::make_shared<Tensor>();
  t->lens = std::move(lens);
  t->complex = complex;
  t->values.assign(size_t(t->size()) * (complex ? 2 : 1), 0.0);
  return t;
}

sisi_engine *engine_of(Algorithm &a, bool complex_path = false) {
  return engine_for(a.getTensorPtr("CoulombVertex"), a.getTensorPtr("HoleEigenEnergies"),
                    a.getTensorPtr("ParticleEigenEnergies"), complex_path);
}

// The reference runs the complex path when the integrals it is given are
// complex tensors (ClusterSinglesDoublesAlgorithm.cxx:25-35).
bool complex_integrals(Algorithm &a) {
  return a.isArgumentGiven("PPHHCoulombIntegrals") &&
         a.getTensorArgument("PPHHCoulombIntegrals").complex;
}

// Amplitudes of the path's type: a real tensor given to the complex path
// gets a zero imaginary part; a complex one given to the real path must
// have none.
std::vector<double> amplitudes_for(const Tensor &t, bool complex_path, const char *name) {
  if (t.complex == complex_path) return t.values;
  const size_t n = size_t(t.size());
  std::vector<double> out(complex_path ? 2 * n : n, 0.0);
  for (size_t k = 0; k < n; ++k) {
    if (complex_path) {
      out[2 * k] = t.values[k];
    } else {
      out[k] = t.values[2 * k];
      if (t.values[2 * k + 1] != 0.0)
        throw StepError(std::string(name) + " is complex but the integrals are real");
    }
  }
  return out;
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
// integrals 203-571, complex: 1 calculateComplexIntegrals 613-810):
// publishes the requested classes (complex tensors with complex: 1).
class CoulombIntegralsFromVertex : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "CoulombIntegralsFromVertex"; }
  void run() override {
    if (getIntegerArgument("antisymmetrize", 0) != 0)
      throw StepError("antisymmetrized integrals are not supported");
    const bool cplx = getIntegerArgument("complex", 0) != 0;
    sisi_engine *e = engine_of(*this, cplx);
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
                   {"HHPPCoulombIntegrals", SISI_HHPP, {o, o, v, v}},
                   {"PHHPCoulombIntegrals", SISI_PHHP, {v, o, o, v}},
                   {"PHHHCoulombIntegrals", SISI_PHHH, {v, o, o, o}},
                   {"PPPPCoulombIntegrals", SISI_PPPP, {v, v, v, v}},
                   {"PPPHCoulombIntegrals", SISI_PPPH, {v, v, v, o}}};
    for (auto &c : classes) {
      if (!out_.count(c.name)) continue;
      if (!cplx && c.cls == SISI_PHHH)
        throw StepError(std::string(c.name) + " is built with complex: 1 only");
      if (cplx && (c.cls == SISI_PPPP || c.cls == SISI_PPPH))
        throw StepError(std::string(c.name) + " is built for real integrals only");
      auto t = make_tensor(c.lens, cplx);
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
    emitted.scalars.push_back({"complex", cplx ? "1" : "0"});
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
    const bool cplx = complex_integrals(*this);
    sisi_engine *e = engine_of(*this, cplx);
    check(sisi_engine_set_formulation(e, SISI_FORMULATION_VERTEX), "sisi_engine_set_formulation");
    solve(e, cplx, true);
  }

 protected:
  // ClusterSinglesDoublesAlgorithm::run on engine e; vertex_arguments: the
  // CoulombFactors / ppl / distinguishable arguments of the vertex residual
  // abbreviation: prefix of the output names (getDataName: "Ccsd", "Drccd",
  // "Uccsd"); kind: 0 restricted CCSD, 1 direct-ring CCD (takes
  // `unrestricted`), 2 spin-orbital CCSD (antisymmetrized energy)
  void solve(sisi_engine *e, bool cplx, bool vertex_arguments, const std::string &abbreviation = "Ccsd",
             int kind = 0) {
    const int64_t o = getTensorArgument("HoleEigenEnergies").lens[0];
    const int64_t v = getTensorArgument("ParticleEigenEnergies").lens[0];
    // getEnergy's arguments (ClusterSinglesDoublesAlgorithm.cxx:133-134,
    // 199-211); the spin-orbital CCSD energy is the antisymmetrized one
    const bool antisymmetrize = getIntegerArgument("antisymmetrize", 0) != 0;
    if (kind != 2)  // (direct-ring CCD's residual reads the same `unrestricted`)
      check(sisi_engine_set_energy_variant(e, getIntegerArgument("unrestricted", 0) != 0,
                                           antisymmetrize),
            "sisi_engine_set_energy_variant");
    if (isArgumentGiven("HPFockMatrix")) {
      if (kind == 2) throw StepError("HPFockMatrix is not read by the spin-orbital CCSD step");
      const auto &F = getTensorArgument("HPFockMatrix");
      if (F.lens != std::vector<int64_t>{o, v})
        throw StepError("HPFockMatrix must be [No, Nv]");
      const auto f = amplitudes_for(F, cplx, "HPFockMatrix");
      check(sisi_engine_set_hp_fock(e, f.data()), "sisi_engine_set_hp_fock");
    } else if (kind != 2) {
      check(sisi_engine_set_hp_fock(e, nullptr), "sisi_engine_set_hp_fock");
    }
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
    opts.bulk_refine_interval = int(getIntegerArgument("bulkRefineInterval", 0));
    if (vertex_arguments) {
      opts.ppl = getIntegerArgument("ppl", 1) != 0;                          // cxx:324
      opts.distinguishable = getIntegerArgument("distinguishable", 0) != 0;  // cxx:123-124
    }
    {
      const std::string rp = getTextArgument("refinePrecision", "bf16");
      if (rp == "bf16" || rp == "3") opts.refine_precision = SISI_BF16;
      else if (rp == "bf16x3" || rp == "2") opts.refine_precision = SISI_BF16X3;
      else if (rp == "fp32" || rp == "1") opts.refine_precision = SISI_FP32;
      else throw StepError("unknown refinePrecision " + rp + " (bf16, bf16x3, fp32)");
    }
    if (vertex_arguments && isArgumentGiven("CoulombFactors")) {
      // factor-based ladder (cxx:340-372): CoulombFactors [NG, NR] and
      // FactorOrbitals [Np, NR], complex or real
      const auto &L = getTensorArgument("CoulombFactors");
      const auto &P = getTensorArgument("FactorOrbitals");
      if (L.lens.size() != 2 || P.lens.size() != 2 || L.lens[1] != P.lens[1] ||
          P.lens[0] != o + v)
        throw StepError("CoulombFactors [NG, NR] / FactorOrbitals [Np, NR] do not match");
      const int64_t NR = L.lens[1];
      int fs = int(getIntegerArgument("factorsSliceSize", -1));
      if (fs == -1 && isArgumentGiven("factorsSliceFactor"))
        fs = int(double(NR) * getRealArgument("factorsSliceFactor"));
      const auto Lz = amplitudes_for(L, true, "CoulombFactors");
      const auto Pz = amplitudes_for(P, true, "FactorOrbitals");
      // complex amplitudes read OutgoingFactorOrbitals as well (cxx:751-752)
      std::vector<double> Oz;
      if (cplx) {
        const auto &O = getTensorArgument("OutgoingFactorOrbitals");
        if (O.lens != P.lens) throw StepError("OutgoingFactorOrbitals [Np, NR] does not match");
        Oz = amplitudes_for(O, true, "OutgoingFactorOrbitals");
      }
      check(sisi_engine_set_coulomb_factors(e, int(NR), fs, Lz.data(), Pz.data(),
                                            Oz.empty() ? nullptr : Oz.data()),
            "sisi_engine_set_coulomb_factors");
    } else {
      check(sisi_engine_set_coulomb_factors(e, 0, 0, nullptr, nullptr, nullptr),
            "sisi_engine_set_coulomb_factors");
    }
    std::vector<double> t1, t2;
    if (isArgumentGiven("initialSinglesAmplitudes"))
      t1 = amplitudes_for(getTensorArgument("initialSinglesAmplitudes"), cplx,
                          "initialSinglesAmplitudes");
    if (isArgumentGiven("initialDoublesAmplitudes"))
      t2 = amplitudes_for(getTensorArgument("initialDoublesAmplitudes"), cplx,
                          "initialDoublesAmplitudes");
    const int64_t w = cplx ? 2 : 1;  // doubles per element
    if (!t1.empty() && int64_t(t1.size()) != w * v * o)
      throw StepError("initialSinglesAmplitudes has the wrong size");
    if (!t2.empty() && int64_t(t2.size()) != w * v * v * o * o)
      throw StepError("initialDoublesAmplitudes has the wrong size");
    if (!t2.empty() && t1.empty()) t1.assign(size_t(w * v * o), 0.0);
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
    // energy decomposition of getEnergy (ClusterSinglesDoublesAlgorithm.cxx:180-190),
    // not emitted by its antisymmetrized branch
    if (isArgumentGiven("HPFockMatrix")) {  // cxx:206-210
      double nc = 0.0;
      check(sisi_engine_get_noncanonical_energy(e, &nc), "sisi_engine_get_noncanonical_energy");
      emitted.scalars.push_back({"noncanonical", num(nc)});
    }
    if (kind != 2 && !antisymmetrize) {
      emitted.scalars.push_back({"direct", num(r.direct)});
      emitted.scalars.push_back({"exchange", num(r.exchange)});
      emitted.scalars.push_back({"singlet", num(0.25 * r.direct - 0.5 * r.exchange)});
      emitted.scalars.push_back({"triplet", num(0.75 * r.direct + 1.5 * r.exchange)});
    }
    if (isArgumentGiven("PairEnergy")) {  // cxx:218-235
      auto P = make_tensor({o, o}, cplx);
      check(sisi_ccsd_pair_energies(e, P->values.data()), "sisi_ccsd_pair_energies");
      allocatedTensorArgument("PairEnergy", P);
    }
    const std::string singles = abbreviation + "SinglesAmplitudes",
                      doubles = abbreviation + "DoublesAmplitudes";
    if (isArgumentGiven(singles) || isArgumentGiven(doubles)) {
      auto T1 = make_tensor({v, o}, cplx), T2 = make_tensor({v, v, o, o}, cplx);
      check(sisi_ccsd_get_amplitudes(e, T1->values.data(), T2->values.data()),
            "sisi_ccsd_get_amplitudes");
      allocatedTensorArgument(singles, T1);
      allocatedTensorArgument(doubles, T2);
    }
    setRealArgument(abbreviation + "Energy", r.energy);
  }
};

// CcsdEnergyFromCoulombIntegralsReference (CcsdEnergyFromCoulombIntegralsReference.cxx:
// 29-295): the same driver on the full-integral residual from the stored
// PPHH, PHPH, HHHH, HHHP, PPPP and PPPH integrals (no vertex); onlyPPL = 1
// keeps only the particle-particle ladder (cxx:56-73).
class CcsdEnergyFromCoulombIntegralsReference : public CcsdEnergyFromCoulombIntegrals {
 public:
  using CcsdEnergyFromCoulombIntegrals::CcsdEnergyFromCoulombIntegrals;
  std::string getName() override { return "CcsdEnergyFromCoulombIntegralsReference"; }
  void run() override {
    const auto &ei = getTensorArgument("HoleEigenEnergies");
    const auto &ea = getTensorArgument("ParticleEigenEnergies");
    if (complex_integrals(*this))
      throw StepError("CcsdEnergyFromCoulombIntegralsReference is built for real integrals");
    const int64_t o = ei.lens[0], v = ea.lens[0];
    sisi_dims d{int(o), int(v), 1};
    struct Owned {
      sisi_engine *e = nullptr;
      ~Owned() { if (e) sisi_engine_destroy(e); }
    } owned;
    check(sisi_engine_create(&d, &owned.e), "sisi_engine_create");
    sisi_engine *e = owned.e;
    check(sisi_engine_set_eigenenergies(e, ei.values.data(), ea.values.data()),
          "sisi_engine_set_eigenenergies");
    const struct {
      const char *name;
      int cls;
      std::vector<int64_t> lens;
    } classes[] = {{"PPHHCoulombIntegrals", SISI_PPHH, {v, v, o, o}},
                   {"PHPHCoulombIntegrals", SISI_PHPH, {v, o, v, o}},
                   {"HHHHCoulombIntegrals", SISI_HHHH, {o, o, o, o}},
                   {"HHHPCoulombIntegrals", SISI_HHHP, {o, o, o, v}},
                   {"PPPPCoulombIntegrals", SISI_PPPP, {v, v, v, v}},
                   {"PPPHCoulombIntegrals", SISI_PPPH, {v, v, v, o}}};
    const bool only_ppl = getIntegerArgument("onlyPPL", 0) == 1;
    for (auto &c : classes) {
      if (!isArgumentGiven(c.name)) {
        if (only_ppl && c.cls != SISI_PPPP) continue;
        throw StepError(std::string(c.name) + " is needed");
      }
      const auto &t = getTensorArgument(c.name);
      if (t.complex || t.lens != c.lens) throw StepError(std::string(c.name) + " has the wrong shape");
      check(sisi_engine_set_integrals(e, c.cls, t.values.data()), "sisi_engine_set_integrals");
    }
    check(sisi_engine_set_formulation(
              e, only_ppl ? SISI_FORMULATION_ONLY_PPL : SISI_FORMULATION_INTEGRALS),
          "sisi_engine_set_formulation");
    solve(e, false, false);
  }
};

// DrccdEnergyFromCoulombIntegrals (DrccdEnergyFromCoulombIntegrals.cxx:31-80)
// on the same driver: the direct-ring CCD residual from the real PPHH, PHHP
// and HHPP integrals, arguments unrestricted, linearized and
// adjacentPairsExchange; outputs DrccdEnergy, DrccdSinglesAmplitudes and
// DrccdDoublesAmplitudes.
class DrccdEnergyFromCoulombIntegrals : public CcsdEnergyFromCoulombIntegrals {
 public:
  using CcsdEnergyFromCoulombIntegrals::CcsdEnergyFromCoulombIntegrals;
  std::string getName() override { return "DrccdEnergyFromCoulombIntegrals"; }
  void run() override {
    const auto &ei = getTensorArgument("HoleEigenEnergies");
    const auto &ea = getTensorArgument("ParticleEigenEnergies");
    if (complex_integrals(*this))
      throw StepError("DrccdEnergyFromCoulombIntegrals is built for real integrals");
    const int64_t o = ei.lens[0], v = ea.lens[0];
    sisi_dims d{int(o), int(v), 1};
    struct Owned {
      sisi_engine *e = nullptr;
      ~Owned() { if (e) sisi_engine_destroy(e); }
    } owned;
    check(sisi_engine_create(&d, &owned.e), "sisi_engine_create");
    sisi_engine *e = owned.e;
    check(sisi_engine_set_eigenenergies(e, ei.values.data(), ea.values.data()),
          "sisi_engine_set_eigenenergies");
    const struct {
      const char *name;
      int cls;
      std::vector<int64_t> lens;
    } classes[] = {{"PPHHCoulombIntegrals", SISI_PPHH, {v, v, o, o}},   // cxx:36-38
                   {"PHHPCoulombIntegrals", SISI_PHHP, {v, o, o, v}},
                   {"HHPPCoulombIntegrals", SISI_HHPP, {o, o, v, v}}};
    for (auto &c : classes) {
      const auto &t = getTensorArgument(c.name);
      if (t.complex || t.lens != c.lens) throw StepError(std::string(c.name) + " has the wrong shape");
      check(sisi_engine_set_integrals(e, c.cls, t.values.data()), "sisi_engine_set_integrals");
    }
    check(sisi_engine_set_drccd(e, int(getIntegerArgument("unrestricted", 0)),
                                int(getIntegerArgument("linearized", 0)),
                                int(getIntegerArgument("adjacentPairsExchange", 0))),
          "sisi_engine_set_drccd");
    check(sisi_engine_set_formulation(e, SISI_FORMULATION_DRCCD), "sisi_engine_set_formulation");
    solve(e, false, false, "Drccd", 1);
  }
};

// UccsdAmplitudesFromCoulombIntegrals (UccsdAmplitudesFromCoulombIntegrals.cxx:
// 107-241): spin-orbital CCSD on the same driver from antisymmetrized real
// integrals, canonical orbitals (HoleEigenEnergies / ParticleEigenEnergies
// of the spin orbitals); reads the PPHH, HHHH, PPPP, HHHP, PPPH and PHPH
// classes (the reference's other classes are these with indices reordered);
// outputs UccsdEnergy, UccsdSinglesAmplitudes, UccsdDoublesAmplitudes.
class UccsdAmplitudesFromCoulombIntegrals : public CcsdEnergyFromCoulombIntegrals {
 public:
  using CcsdEnergyFromCoulombIntegrals::CcsdEnergyFromCoulombIntegrals;
  std::string getName() override { return "UccsdAmplitudesFromCoulombIntegrals"; }
  void run() override {
    if (getIntegerArgument("onlyPPL", 0) == 1) throw StepError("onlyPPL is not supported for UCCSD");
    if (isArgumentGiven("HPFockMatrix") || isArgumentGiven("HHFockMatrix") ||
        isArgumentGiven("PPFockMatrix"))
      throw StepError("non-canonical orbitals (Fock matrices) are not supported");
    const auto &ei = getTensorArgument("HoleEigenEnergies");
    const auto &ea = getTensorArgument("ParticleEigenEnergies");
    if (complex_integrals(*this))
      throw StepError("UccsdAmplitudesFromCoulombIntegrals is built for real integrals");
    const int64_t o = ei.lens[0], v = ea.lens[0];
    sisi_dims d{int(o), int(v), 1};
    struct Owned {
      sisi_engine *e = nullptr;
      ~Owned() { if (e) sisi_engine_destroy(e); }
    } owned;
    check(sisi_engine_create(&d, &owned.e), "sisi_engine_create");
    sisi_engine *e = owned.e;
    check(sisi_engine_set_eigenenergies(e, ei.values.data(), ea.values.data()),
          "sisi_engine_set_eigenenergies");
    const struct {
      const char *name;
      int cls;
      std::vector<int64_t> lens;
    } classes[] = {{"PPHHCoulombIntegrals", SISI_PPHH, {v, v, o, o}},   // cxx:138, 167-180
                   {"HHHHCoulombIntegrals", SISI_HHHH, {o, o, o, o}},
                   {"PPPPCoulombIntegrals", SISI_PPPP, {v, v, v, v}},
                   {"HHHPCoulombIntegrals", SISI_HHHP, {o, o, o, v}},
                   {"PPPHCoulombIntegrals", SISI_PPPH, {v, v, v, o}},
                   {"PHPHCoulombIntegrals", SISI_PHPH, {v, o, v, o}}};
    for (auto &c : classes) {
      const auto &t = getTensorArgument(c.name);
      if (t.complex || t.lens != c.lens) throw StepError(std::string(c.name) + " has the wrong shape");
      check(sisi_engine_set_integrals(e, c.cls, t.values.data()), "sisi_engine_set_integrals");
    }
    check(sisi_engine_set_formulation(e, SISI_FORMULATION_UCCSD), "sisi_engine_set_formulation");
    solve(e, false, false, "Uccsd", 2);
  }
};

// CcsdPerturbativeTriples (CcsdPerturbativeTriples.cxx:119-248): output is
// CCSD energy + (T).
class CcsdPerturbativeTriples : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "CcsdPerturbativeTriples"; }
  void run() override {
    if (complex_integrals(*this))
      throw StepError("complex integrals: (T) of the complex path is CcsdPerturbativeTriplesComplex");
    sisi_engine *e = engine_of(*this);
    const auto &T1 = getTensorArgument("CcsdSinglesAmplitudes");
    const auto &T2 = getTensorArgument("CcsdDoublesAmplitudes");
    if (T1.complex || T2.complex) throw StepError("CcsdPerturbativeTriples needs real amplitudes");
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

// CcsdPerturbativeTriplesComplex (CcsdPerturbativeTriplesComplex.cxx:32-85):
// the complex calculator when the PPHH integrals are complex, else the real
// one; output is CCSD energy + Re (T).
class CcsdPerturbativeTriplesComplex : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "CcsdPerturbativeTriplesComplex"; }
  void run() override {
    const bool cplx = complex_integrals(*this);
    sisi_engine *e = engine_of(*this, cplx);
    const int64_t o = getTensorArgument("HoleEigenEnergies").lens[0];
    const int64_t v = getTensorArgument("ParticleEigenEnergies").lens[0];
    const auto T1 = amplitudes_for(getTensorArgument("CcsdSinglesAmplitudes"), cplx,
                                   "CcsdSinglesAmplitudes");
    const auto T2 = amplitudes_for(getTensorArgument("CcsdDoublesAmplitudes"), cplx,
                                   "CcsdDoublesAmplitudes");
    const int64_t w = cplx ? 2 : 1;
    if (int64_t(T1.size()) != w * v * o || int64_t(T2.size()) != w * v * v * o * o)
      throw StepError("amplitudes do not match the eigenenergies");
    double et = 0;
    check(sisi_triples_energy(e, T1.data(), T2.data(), precision_argument(*this), &et),
          "sisi_triples_energy");
    const double ec = getRealArgument("CcsdEnergy");
    emitted.scalars.push_back({"triples", num(et)});
    emitted.scalars.push_back({"ccsd", num(ec)});
    emitted.scalars.push_back({"energy", num(ec + et)});
    setRealArgument("CcsdPerturbativeTriplesComplexEnergy", ec + et);
  }
};

// UCcsdIPEquationOfMotionDavidson (UCcsdIPEquationOfMotionDavidson.cxx:25-206)
// and UCcsdEAEquationOfMotionDavidson (UCcsdEAEquationOfMotionDavidson.cxx:
// 25-351): the lowest IP- / EA-EOM-CCSD roots of spin-orbital CCSD amplitudes
// by the host Davidson driver (EigenSystemDavidsonMono), each right-apply on
// the device (sisi_eom_ip_sigma / sisi_eom_ea_sigma).  Inputs as the
// reference's: the eigenenergies, SinglesAmplitudes [Nv,No],
// DoublesAmplitudes [Nv,Nv,No,No] and antisymmetrized classes; IP reads
// HHPP, HHHH, HHHP, HPPP and HPHP, EA reads HHPP, PPPP, HHHP, PPPH and PHPH
// (the other classes the reference reads are these with indices reordered).
// Vector spaces: IP r1[i] and r2[i,j,a] for i < j, EA r1[a] and r2[j,a,b]
// for a < b; preconditioner diagonals -f_ii, -f_ii - f_jj + f_aa (IP) and
// f_aa, -f_jj + f_aa + f_bb (EA) as IPCcsdPreconditioner's and
// EACcsdPreconditioner's (CcsdPreconditioner.cxx:546-560, 374-397).  Emits the
// eigenvalues and writes IPEomEnergies.dat / EAEomEnergies.dat ("k re im" per
// root) as the reference does (IP cxx:199-205, EA cxx:343-350).
class EomCcsdDavidson : public Algorithm {
 public:
  EomCcsdDavidson(std::map<std::string, Argument> in, std::map<std::string, Argument> out, DataMap &d,
                  bool ea)
      : Algorithm(std::move(in), std::move(out), d), ea_(ea) {}
  void run() override {
    if (isArgumentGiven("HPFockMatrix") || isArgumentGiven("HHFockMatrix") ||
        isArgumentGiven("PPFockMatrix"))
      throw StepError("non-canonical orbitals (Fock matrices) are not supported");
    for (const char *unsupported : {"refreshIterations", "oneBodyRdmRange", "printEigenvectorsRange"})
      if (!getTextArgument(unsupported, "").empty())
        throw StepError(std::string(unsupported) + " is not supported");
    const auto &ei = getTensorArgument("HoleEigenEnergies");
    const auto &ea = getTensorArgument("ParticleEigenEnergies");
    const int64_t o = ei.lens[0], v = ea.lens[0];
    const auto &T1 = getTensorArgument("SinglesAmplitudes");
    const auto &T2 = getTensorArgument("DoublesAmplitudes");
    if (T1.complex || T2.complex || T1.lens != std::vector<int64_t>{v, o} ||
        T2.lens != std::vector<int64_t>{v, v, o, o})
      throw StepError("the amplitudes must be real with shapes [Nv,No] and [Nv,Nv,No,No]");
    DavidsonOptions opt;
    opt.nroots = int(getIntegerArgument("eigenstates", 1));
    opt.tolerance_energy = getRealArgument("energyConvergence", 1e-6);
    opt.tolerance_vectors = getRealArgument("amplitudesConvergence", 1e-6);
    opt.max_iterations = int(getIntegerArgument("maxIterations", 32));
    opt.min_iterations = int(getIntegerArgument("minIterations", 1));
    opt.max_basis = int(std::min<int64_t>(
        getIntegerArgument("maxBasisSize", o * v + (o * (o - 1) / 2) * (v * (v - 1) / 2)), INT32_MAX));
    opt.refresh_on_max_basis = getIntegerArgument("refreshOnMaxBasisSize", 0) == 1;

    sisi_dims d{int(o), int(v), 1};
    struct Owned {
      sisi_engine *e = nullptr;
      ~Owned() { if (e) sisi_engine_destroy(e); }
    } owned;
    check(sisi_engine_create(&d, &owned.e), "sisi_engine_create");
    sisi_engine *e = owned.e;
    check(sisi_engine_set_eigenenergies(e, ei.values.data(), ea.values.data()),
          "sisi_engine_set_eigenenergies");
    upload_classes(e, o, v);
    check(sisi_engine_set_formulation(e, SISI_FORMULATION_UCCSD), "sisi_engine_set_formulation");

    const EomSpace space(ea_, o, v);
    if (opt.nroots < 1 || size_t(opt.nroots) > space.n)
      throw StepError("eigenstates must lie in [1, dimension]");
    const std::vector<double> diag = space.diagonal(ei.values.data(), ea.values.data());
    std::vector<double> r1, r2, s1(space.r1_size()), s2(space.r2_size());
    auto apply = [&](const std::vector<double> &xv) {
      space.unpack(xv, r1, r2);
      check((ea_ ? sisi_eom_ea_sigma : sisi_eom_ip_sigma)(e, T1.values.data(), T2.values.data(), r1.data(),
                                                         r2.data(), s1.data(), s2.data()),
            ea_ ? "sisi_eom_ea_sigma" : "sisi_eom_ip_sigma");
      return space.pack(s1, s2);
    };
    const DavidsonResult res = davidson(apply, diag, lowest_diagonal_basis(diag, opt.nroots), opt);
    std::ofstream dat(ea_ ? "EAEomEnergies.dat" : "IPEomEnergies.dat");
    for (size_t k = 0; k < res.values.size(); ++k) {
      dat << (k + 1) << " " << num(res.values[k].real()) << " " << num(res.values[k].imag()) << "\n";
      emitted.scalars.push_back({"eigenvalue" + std::to_string(k + 1), num(res.values[k].real())});
    }
    emitted.scalars.push_back({"davidsonIterations", std::to_string(res.iterations)});
    emitted.scalars.push_back({"converged", res.converged ? "true" : "false"});
  }

 private:
  bool ea_;

  const Tensor &real_class(const char *name, std::vector<int64_t> lens) {
    const auto &t = getTensorArgument(name);
    if (t.complex || t.lens != lens) throw StepError(std::string(name) + " has the wrong shape");
    return t;
  }

  // reference classes -> engine classes (pphh_from_hhpp, ppph_from_hppp,
  // phph_from_hphp); EA's PPPH and PHPH come in the engine's order
  void upload_classes(sisi_engine *e, int64_t o, int64_t v) {
    auto set = [&](int cls, const double *p) { check(sisi_engine_set_integrals(e, cls, p), "sisi_engine_set_integrals"); };
    set(SISI_PPHH, pphh_from_hhpp(real_class("HHPPCoulombIntegrals", {o, o, v, v}).values.data(), o, v).data());
    set(SISI_HHHP, real_class("HHHPCoulombIntegrals", {o, o, o, v}).values.data());
    if (ea_) {
      set(SISI_PPPP, real_class("PPPPCoulombIntegrals", {v, v, v, v}).values.data());
      set(SISI_PPPH, real_class("PPPHCoulombIntegrals", {v, v, v, o}).values.data());
      set(SISI_PHPH, real_class("PHPHCoulombIntegrals", {v, o, v, o}).values.data());
      return;
    }
    set(SISI_HHHH, real_class("HHHHCoulombIntegrals", {o, o, o, o}).values.data());
    set(SISI_PPPH, ppph_from_hppp(real_class("HPPPCoulombIntegrals", {o, v, v, v}).values.data(), o, v).data());
    set(SISI_PHPH, phph_from_hphp(real_class("HPHPCoulombIntegrals", {o, v, o, v}).values.data(), o, v).data());
  }
};

class UCcsdIPEquationOfMotionDavidson : public EomCcsdDavidson {
 public:
  UCcsdIPEquationOfMotionDavidson(std::map<std::string, Argument> in, std::map<std::string, Argument> out,
                                  DataMap &d)
      : EomCcsdDavidson(std::move(in), std::move(out), d, false) {}
  std::string getName() override { return "UCcsdIPEquationOfMotionDavidson"; }
};

class UCcsdEAEquationOfMotionDavidson : public EomCcsdDavidson {
 public:
  UCcsdEAEquationOfMotionDavidson(std::map<std::string, Argument> in, std::map<std::string, Argument> out,
                                  DataMap &d)
      : EomCcsdDavidson(std::move(in), std::move(out), d, true) {}
  std::string getName() override { return "UCcsdEAEquationOfMotionDavidson"; }
};

static Registrar<GenerateSyntheticSystem> r0("GenerateSyntheticSystem");
static Registrar<CcsdPerturbativeTriplesComplex> r5("CcsdPerturbativeTriplesComplex");
static Registrar<CcsdEnergyFromCoulombIntegralsReference> r6("CcsdEnergyFromCoulombIntegralsReference");
static Registrar<CoulombIntegralsFromVertex> r1("CoulombIntegralsFromVertex");
static Registrar<UegVertexGenerator> r4("UegVertexGenerator");
static Registrar<CcsdEnergyFromCoulombIntegrals> r2("CcsdEnergyFromCoulombIntegrals");
static Registrar<CcsdPerturbativeTriples> r3("CcsdPerturbativeTriples");
static Registrar<DrccdEnergyFromCoulombIntegrals> r7("DrccdEnergyFromCoulombIntegrals");
static Registrar<UccsdAmplitudesFromCoulombIntegrals> r8("UccsdAmplitudesFromCoulombIntegrals");
static Registrar<UCcsdIPEquationOfMotionDavidson> r9("UCcsdIPEquationOfMotionDavidson");
static Registrar<UCcsdEAEquationOfMotionDavidson> r10("UCcsdEAEquationOfMotionDavidson");

}  // namespace sisi_host
