// Input readers on the caller's side of the path: the Coulomb vertex and
// eigenenergies from the files the reference reads, and TensorNorm for
// checking what was read.  A complex vertex with a non-zero imaginary part
// is kept complex (for the complex path, complex: 1); one whose imaginary
// part is zero is stored real.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <map>
#include <set>
#include <fstream>
#include <sstream>

#include "algorithm.hpp"
#include "yaml_lite.hpp"

namespace sisi_host {

namespace {

std::shared_ptr<Tensor> new_tensor(std::vector<int64_t> lens) {
  auto t = std::make_shared<Tensor>();
  t->lens = std::move(lens);
  t->values.assign(size_t(t->size()), 0.0);
  return t;
}

std::string num(double v) {
  std::ostringstream o;
  o.precision(17);
  o << v;
  return o.str();
}

const YamlNode &require(const YamlNode &n, const std::string &key, const std::string &file) {
  const YamlNode *c = n.type == YamlNode::Map ? n.get(key) : nullptr;
  if (!c) throw StepError(file + ": missing key '" + key + "'");
  return *c;
}

double to_double(const YamlNode &n, const std::string &what) {
  if (n.type != YamlNode::Scalar) throw StepError(what + ": expected a number");
  char *end = nullptr;
  const double v = std::strtod(n.scalar.c_str(), &end);
  if (!end || *end || n.scalar.empty()) throw StepError(what + ": '" + n.scalar + "' is not a number");
  return v;
}

}  // namespace

// CoulombVertexReader (CoulombVertexReader.cxx:26-111, format in
// docs/manual.org "FTODDUMP"): a 32-byte header {magic[8], int32 No, Nv, NG,
// NSpins, kPoints, reserved} and chunks {magic[8], int64 size of the whole
// chunk, payload}: FTODreal / FTODimag hold Gamma[G,p,q] (G fastest),
// FTODepsi holds eps_p (holes first); other chunks are skipped by size.
// The reference compares the first 8 bytes of "sisi4sFTOD"; the documented
// cc4s magic "cc4sFTOD" is accepted as well.
class CoulombVertexReader : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "CoulombVertexReader"; }
  void run() override {
    const std::string fileName = getTextArgument("file");
    if (getIntegerArgument("unrestricted", 0) != 0)
      throw StepError("CoulombVertexReader: unrestricted = 1 is not supported");
    std::ifstream f(fileName, std::ios::binary);
    if (!f) throw StepError("Failed to open file " + fileName);
    f.seekg(0, std::ios::end);
    const int64_t fileSize = int64_t(f.tellg());
    f.seekg(0);
    struct {
      char magic[8];
      int32_t No, Nv, NG, NSpins, kPoints, reserved;
    } header;
    static_assert(sizeof(header) == 32, "FTOD header is 32 bytes");
    if (fileSize < int64_t(sizeof(header)) || !f.read(reinterpret_cast<char *>(&header), sizeof(header)))
      throw StepError("Invalid file format: " + fileName + " is shorter than its header");
    if (std::strncmp(header.magic, "sisi4sFT", 8) != 0 && std::strncmp(header.magic, "cc4sFTOD", 8) != 0)
      throw StepError("Invalid file format: " + fileName);
    const int64_t No = header.No, Nv = header.Nv, NG = header.NG, Np = No + Nv;
    if (No <= 0 || Nv <= 0 || NG <= 0) throw StepError("Invalid file format: bad dimensions");
    auto gamma = new_tensor({NG, Np, Np});
    auto epsi = new_tensor({No}), epsa = new_tensor({Nv});
    bool have_real = false, have_eps = false;
    double imag_max = 0.0;
    const int64_t nvert = NG * Np * Np;
    std::vector<double> imag;
    int64_t offset = sizeof(header);
    while (offset < fileSize) {
      struct {
        char magic[8];
        int64_t size;
      } chunk;
      static_assert(sizeof(chunk) == 16, "FTOD chunk header is 16 bytes");
      f.seekg(offset);
      if (!f.read(reinterpret_cast<char *>(&chunk), sizeof(chunk)) || chunk.size < 16 ||
          offset + chunk.size > fileSize)
        throw StepError("Invalid file format: truncated chunk at offset " + std::to_string(offset));
      const int64_t payload = chunk.size - 16;
      if (std::strncmp(chunk.magic, "FTODreal", 8) == 0 ||
          std::strncmp(chunk.magic, "FTODimag", 8) == 0) {
        if (payload < 8 * nvert) throw StepError("Invalid file format: short vertex chunk");
        const bool real = chunk.magic[4] == 'r';
        double *dst = gamma->values.data();
        if (!real) {
          imag.resize(size_t(nvert));
          dst = imag.data();
        }
        if (!f.read(reinterpret_cast<char *>(dst), 8 * nvert))
          throw StepError("Invalid file format: short vertex chunk");
        if (real) have_real = true;
        else
          for (double x : imag) imag_max = std::max(imag_max, std::fabs(x));
      } else if (std::strncmp(chunk.magic, "FTODepsi", 8) == 0) {
        if (payload < 8 * Np) throw StepError("Invalid file format: short eigenenergy chunk");
        if (!f.read(reinterpret_cast<char *>(epsi->values.data()), 8 * No) ||
            !f.read(reinterpret_cast<char *>(epsa->values.data()), 8 * Nv))
          throw StepError("Invalid file format: short eigenenergy chunk");
        have_eps = true;
      }
      offset += chunk.size;
    }
    if (!have_real || !have_eps)
      throw StepError("Invalid file format: " + fileName + " lacks the FTODreal or FTODepsi chunk");
    if (imag_max != 0.0) {  // complex vertex: interleave (re, im)
      auto z = std::make_shared<Tensor>();
      z->lens = gamma->lens;
      z->complex = true;
      z->values.resize(size_t(2 * nvert));
      for (int64_t k = 0; k < nvert; ++k) {
        z->values[size_t(2 * k)] = gamma->values[size_t(k)];
        z->values[size_t(2 * k + 1)] = imag[size_t(k)];
      }
      gamma = z;
    }
    allocatedTensorArgument("CoulombVertex", gamma);
    allocatedTensorArgument("HoleEigenEnergies", epsi);
    allocatedTensorArgument("ParticleEigenEnergies", epsa);
    emitted.scalars.push_back({"NG", std::to_string(NG)});
    emitted.scalars.push_back({"No", std::to_string(No)});
    emitted.scalars.push_back({"Nv", std::to_string(Nv)});
  }
};

// Read (Read.cxx:24-92): a cc4s tensor header (version 100, type Tensor,
// dimensions [{length, type}], scalarType Real64 | Complex64,
// elements.type TextFile | IeeeBinaryFile, unit) with its values in the file
// of the same name and extension ".elements", first index fastest: one value
// per line (text) or raw doubles ((re, im) pairs for Complex64).
class Read : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "Read"; }
  void run() override {
    const std::string fileName = getTextArgument("fileName");
    YamlNode h;
    try {
      h = parse_yaml_file(fileName);
    } catch (const YamlError &e) {
      throw StepError("Read: " + std::string(e.what()));
    }
    const YamlNode &version = require(h, "version", fileName);
    if (version.type != YamlNode::Scalar || version.scalar != "100")
      throw StepError(fileName + ": unsupported version");
    const YamlNode &type = require(h, "type", fileName);
    if (type.scalar != "Tensor") throw StepError(fileName + ": type must be Tensor");
    require(h, "unit", fileName);
    const std::string scalarType = require(h, "scalarType", fileName).scalar;
    if (scalarType != "Real64" && scalarType != "Complex64")
      throw StepError(fileName + ": scalarType must be Real64 or Complex64");
    const std::string elements = require(require(h, "elements", fileName), "type", fileName).scalar;
    if (elements != "TextFile" && elements != "IeeeBinaryFile")
      throw StepError(fileName + ": elements type must be TextFile or IeeeBinaryFile");
    const YamlNode &dims = require(h, "dimensions", fileName);
    if (dims.type != YamlNode::Seq || dims.seq.empty())
      throw StepError(fileName + ": dimensions must be a non-empty list");
    std::vector<int64_t> lens;
    for (auto &d : dims.seq) {
      const double len = to_double(require(d, "length", fileName), fileName + " dimension length");
      const std::string dt = require(d, "type", fileName).scalar;
      if (dt != "AuxiliaryField" && dt != "State")
        throw StepError(fileName + ": dimension type must be AuxiliaryField or State");
      if (len < 1 || len != std::floor(len)) throw StepError(fileName + ": bad dimension length");
      lens.push_back(int64_t(len));
    }
    std::string data = fileName;
    const auto slash = data.find_last_of('/');
    const auto dot = data.find_last_of('.');
    if (dot != std::string::npos && (slash == std::string::npos || dot > slash)) data.resize(dot);
    data += ".elements";
    auto t = new_tensor(lens);
    const int64_t count = t->size();
    if (scalarType == "Real64" && elements == "TextFile") {
      std::ifstream s(data);
      if (!s) throw StepError("Read: cannot open " + data);
      int64_t k = 0;
      for (std::string line; std::getline(s, line);) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        if (k >= count) throw StepError("Read: " + data + " holds more than " + std::to_string(count) + " values");
        t->values[size_t(k++)] = std::atof(line.c_str());
      }
      if (k != count) throw StepError("Read: " + data + " holds " + std::to_string(k) + " of " + std::to_string(count) + " values");
    } else if (elements == "IeeeBinaryFile") {
      std::ifstream s(data, std::ios::binary);
      if (!s) throw StepError("Read: cannot open " + data);
      if (scalarType == "Real64") {
        if (!s.read(reinterpret_cast<char *>(t->values.data()), 8 * count))
          throw StepError("Read: " + data + " is shorter than " + std::to_string(count) + " doubles");
      } else {
        std::vector<double> z(size_t(2 * count));
        if (!s.read(reinterpret_cast<char *>(z.data()), 16 * count))
          throw StepError("Read: " + data + " is shorter than " + std::to_string(count) + " complex values");
        double imag_max = 0.0;
        for (int64_t k = 0; k < count; ++k) {
          t->values[size_t(k)] = z[size_t(2 * k)];
          imag_max = std::max(imag_max, std::fabs(z[size_t(2 * k + 1)]));
        }
        if (imag_max != 0.0) {  // keep it complex, (re, im) interleaved
          t->complex = true;
          t->values = std::move(z);
        }
      }
    } else {
      throw StepError("Read: Complex64 tensors are read from IeeeBinaryFile elements only");
    }
    allocatedTensorArgument("destination", t);
  }
};

// DefineHolesAndParticles (cc4s/DefineHolesAndParticles.cxx:8-51):
// metaData.energies split at metaData.fermiEnergy (e <= fermiEnergy are
// holes), keeping the file's order within each part.
class DefineHolesAndParticles : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "DefineHolesAndParticles"; }
  void run() override {
    const std::string fileName = getTextArgument("fileName");
    YamlNode h;
    try {
      h = parse_yaml_file(fileName);
    } catch (const YamlError &e) {
      throw StepError("DefineHolesAndParticles: " + std::string(e.what()));
    }
    const YamlNode &meta = require(h, "metaData", fileName);
    const double fermi = to_double(require(meta, "fermiEnergy", fileName), "fermiEnergy");
    const YamlNode &en = require(meta, "energies", fileName);
    if (en.type != YamlNode::Seq) throw StepError(fileName + ": energies must be a list");
    std::vector<double> holes, particles;
    for (auto &e : en.seq) {
      const double v = to_double(e, "energies");
      (v <= fermi ? holes : particles).push_back(v);
    }
    if (holes.empty() || particles.empty())
      throw StepError(fileName + ": no holes or no particles at this Fermi energy");
    auto epsi = new_tensor({int64_t(holes.size())}), epsa = new_tensor({int64_t(particles.size())});
    epsi->values = holes;
    epsa->values = particles;
    allocatedTensorArgument("HoleEigenEnergies", epsi);
    allocatedTensorArgument("ParticleEigenEnergies", epsa);
    emitted.scalars.push_back({"No", std::to_string(holes.size())});
    emitted.scalars.push_back({"Nv", std::to_string(particles.size())});
  }
};

// TensorNorm (TensorNorm.cxx): Frobenius norm of Data, published as Norm.
class TensorNorm : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "TensorNorm"; }
  void run() override {
    const Tensor &t = getTensorArgument("Data");
    long double s = 0.0L;
    for (double x : t.values) s += (long double)x * x;
    const double norm = std::sqrt(double(s));
    emitted.scalars.push_back({"norm", num(norm)});
    if (isArgumentGiven("Norm")) setRealArgument("Norm", norm);
  }
};

// TensorUnrestricter (TensorUnrestricter.cxx:15-79): spatial orbitals to
// spin orbitals, spin orbital 2 p + s.  Order 1 and 2 tensors fill the
// diagonal spin blocks; order 4 tensors <pq|rs> the blocks where indices 0
// and 2, and 1 and 3, carry the same spin.  Real or complex.
class TensorUnrestricter : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "TensorUnrestricter"; }
  void run() override {
    const Tensor &t = getTensorArgument("Data");
    const size_t order = t.lens.size();
    if (order != 1 && order != 2 && order != 4) throw StepError("Tensor size not supported");
    auto out = std::make_shared<Tensor>();
    out->complex = t.complex;
    for (auto l : t.lens) out->lens.push_back(2 * l);
    const int w = t.complex ? 2 : 1;
    out->values.assign(size_t(out->size()) * w, 0.0);
    std::vector<int64_t> idx(order, 0);
    const int64_t n = t.size();
    for (int64_t e = 0; e < n; ++e) {
      int64_t r = e;
      for (size_t k = 0; k < order; ++k) {
        idx[k] = r % t.lens[k];
        r /= t.lens[k];
      }
      for (int s = 0; s < 2; ++s)
        for (int u = 0; u < (order == 4 ? 2 : 1); ++u) {
          // spins of the indices: order 1/2 all s; order 4 (s, u, s, u)
          int64_t o = 0, stride = 1;
          for (size_t k = 0; k < order; ++k) {
            const int spin = (order == 4 && (k == 1 || k == 3)) ? u : s;
            o += (2 * idx[k] + spin) * stride;
            stride *= out->lens[k];
          }
          for (int c = 0; c < w; ++c) out->values[size_t(o * w + c)] = t.values[size_t(e * w + c)];
        }
    }
    allocatedTensorArgument("Out", out);
  }
};

// TensorAntisymmetrizer (TensorAntisymmetrizer.cxx:14-124, Integrals.hpp:
// 52-84): every given class V_pqrs becomes V_pqrs - V_pqsr, the exchange
// value read from the first given class among the equivalent forms of
// <qp|rs> and <pq|sr> (the class itself first), always from the values
// before antisymmetrisation.
class TensorAntisymmetrizer : public Algorithm {
 public:
  using Algorithm::Algorithm;
  std::string getName() override { return "TensorAntisymmetrizer"; }
  void run() override {
    using Info = std::pair<std::string, std::string>;  // class letters, index labels
    static const Info classes[] = {
        {"HHHH", "ijkl"}, {"HHHP", "ijka"}, {"HHPH", "ijak"}, {"HHPP", "ijab"},
        {"HPHH", "iajk"}, {"HPHP", "iajb"}, {"HPPH", "iabj"}, {"HPPP", "iabc"},
        {"PHHH", "aijk"}, {"PHHP", "aijb"}, {"PHPH", "aibj"}, {"PHPP", "aibc"},
        {"PPHH", "abij"}, {"PPHP", "abic"}, {"PPPH", "abci"}, {"PPPP", "abcd"}};
    auto swap2 = [](Info x, int i, int j) {
      std::swap(x.first[size_t(i)], x.first[size_t(j)]);
      std::swap(x.second[size_t(i)], x.second[size_t(j)]);
      return x;
    };
    auto antisymmetrizers = [&](const Info &x) {
      std::vector<Info> res;
      std::set<std::string> names;
      for (const Info &t : {swap2(x, 0, 1), swap2(x, 2, 3)})
        for (const Info &eq : {swap2(swap2(t, 0, 1), 2, 3), swap2(swap2(t, 0, 2), 1, 3),
                               swap2(t, 0, 2)})
          if (names.insert(eq.first).second) res.push_back(eq);
      std::stable_partition(res.begin(), res.end(),
                            [&](const Info &a) { return a.first == x.first; });
      return res;
    };
    std::map<std::string, Tensor> original;
    for (const auto &c : classes)
      if (isArgumentGiven(c.first + "CoulombIntegrals"))
        original[c.first] = getTensorArgument(c.first + "CoulombIntegrals");
    if (original.empty()) throw StepError("TensorAntisymmetrizer: no integrals given");
    for (const auto &c : classes) {
      if (!original.count(c.first)) continue;
      const Info *use = nullptr;
      const auto cands = antisymmetrizers(c);
      for (const auto &a : cands)
        if (original.count(a.first)) {
          use = &a;
          break;
        }
      if (!use) throw StepError(c.first + "CoulombIntegrals could not be antisymmetrized");
      Tensor &V = *getTensorPtr(c.first + "CoulombIntegrals");
      const Tensor &A = original.at(use->first);
      if (V.lens.size() != 4 || A.lens.size() != 4 || V.complex != A.complex)
        throw StepError("TensorAntisymmetrizer: integrals of different kinds");
      // position in A of each of V's index labels
      int pos[4];
      for (int k = 0; k < 4; ++k) {
        const auto at = use->second.find(c.second[size_t(k)]);
        if (at == std::string::npos) throw StepError("TensorAntisymmetrizer: label mismatch");
        pos[k] = int(at);
      }
      int64_t astride[4], s = 1;
      for (int k = 0; k < 4; ++k) {
        astride[k] = s;
        s *= A.lens[size_t(k)];
      }
      for (int k = 0; k < 4; ++k)
        if (A.lens[size_t(pos[k])] != V.lens[size_t(k)])
          throw StepError("TensorAntisymmetrizer: " + use->first + " does not match " + c.first);
      const int w = V.complex ? 2 : 1;
      const int64_t n = V.size();
      for (int64_t e = 0; e < n; ++e) {
        int64_t r = e, ao = 0;
        for (int k = 0; k < 4; ++k) {
          ao += (r % V.lens[size_t(k)]) * astride[pos[k]];
          r /= V.lens[size_t(k)];
        }
        for (int cc = 0; cc < w; ++cc) V.values[size_t(e * w + cc)] -= A.values[size_t(ao * w + cc)];
      }
    }
  }
};

static Registrar<TensorUnrestricter> rr4("TensorUnrestricter");
static Registrar<TensorAntisymmetrizer> rr5("TensorAntisymmetrizer");
static Registrar<CoulombVertexReader> rr0("CoulombVertexReader");
static Registrar<Read> rr1("Read");
static Registrar<DefineHolesAndParticles> rr2("DefineHolesAndParticles");
static Registrar<TensorNorm> rr3("TensorNorm");

}  // namespace sisi_host
