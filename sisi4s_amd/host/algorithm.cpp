#include "algorithm.hpp"

#include <algorithm>
#include <cmath>

#include <cstdlib>

namespace sisi_host {

std::map<std::string, Factory> &registry() {
  static std::map<std::string, Factory> r;
  return r;
}

void check(int rc, const char *what) {
  if (rc != SISI_OK)
    throw StepError(std::string(what) + " failed (" + std::to_string(rc) + "): " +
                    sisi_last_error());
}

const Argument &Algorithm::input(const std::string &n) const {
  auto it = in_.find(n);
  if (it == in_.end()) throw StepError("Missing argument: " + n);
  return it->second;
}

// Algorithm::getTensorArgument (Algorithm.cxx:190-201): the argument must
// name a symbol holding a tensor.
std::shared_ptr<Tensor> Algorithm::getTensorPtr(const std::string &n) {
  const Argument &a = input(n);
  if (!a.symbol) throw StepError("Argument " + n + " must be a $symbol of a tensor");
  Datum *d = data_.find(a.value);
  if (!d || d->type != Datum::TensorT)
    throw StepError("Incompatible type for argument " + n + ": " + a.value + " is not a tensor");
  return d->tensor;
}

static Datum resolve(DataMap &data, const Argument &a, const std::string &n) {
  if (!a.symbol) {
    Datum d;
    d.type = Datum::Text;
    d.text = a.value;
    return d;
  }
  Datum *d = data.find(a.value);
  if (!d) throw StepError("Argument " + n + ": symbol " + a.value + " has no data");
  return *d;
}

// getRealArgument (Algorithm.cxx:100-118): reals, integers, or text that
// parses as a number.
double Algorithm::getRealArgument(const std::string &n) {
  Datum d = resolve(data_, input(n), n);
  if (d.type == Datum::Real) return d.real;
  if (d.type == Datum::Integer) return double(d.integer);
  if (d.type == Datum::Text) {
    char *end = nullptr;
    const double v = std::strtod(d.text.c_str(), &end);
    if (end && *end == 0 && !d.text.empty()) return v;
  }
  throw StepError("Incompatible type for argument " + n + ": expected a real");
}

double Algorithm::getRealArgument(const std::string &n, double def) {
  return in_.count(n) ? getRealArgument(n) : def;
}

int64_t Algorithm::getIntegerArgument(const std::string &n) {
  Datum d = resolve(data_, input(n), n);
  if (d.type == Datum::Integer) return d.integer;
  if (d.type == Datum::Text) {
    char *end = nullptr;
    const long long v = std::strtoll(d.text.c_str(), &end, 10);
    if (end && *end == 0 && !d.text.empty()) return v;
  }
  throw StepError("Incompatible type for argument " + n + ": expected an integer");
}

int64_t Algorithm::getIntegerArgument(const std::string &n, int64_t def) {
  return in_.count(n) ? getIntegerArgument(n) : def;
}

std::string Algorithm::getTextArgument(const std::string &n) {
  Datum d = resolve(data_, input(n), n);
  if (d.type == Datum::Text) return d.text;
  throw StepError("Incompatible type for argument " + n + ": expected text");
}

std::string Algorithm::getTextArgument(const std::string &n, const std::string &def) {
  return in_.count(n) ? getTextArgument(n) : def;
}

// setRealArgument (Algorithm.cxx:364-367): writes to the output symbol.
void Algorithm::setRealArgument(const std::string &n, double v) {
  auto it = out_.find(n);
  if (it == out_.end()) return;  // output not requested
  if (!it->second.symbol) throw StepError("Output " + n + " must be a $symbol");
  Datum d;
  d.type = Datum::Real;
  d.real = v;
  data_.set(it->second.value, d);
}

// allocatedTensorArgument (Algorithm.cxx:289-295): ownership passes to the
// data map, replacing the symbol's previous data.
void Algorithm::allocatedTensorArgument(const std::string &n, std::shared_ptr<Tensor> t) {
  auto it = out_.find(n);
  if (it == out_.end()) return;
  if (!it->second.symbol) throw StepError("Output " + n + " must be a $symbol");
  Datum d;
  d.type = Datum::TensorT;
  d.tensor = std::move(t);
  data_.set(it->second.value, d);
}

namespace {
struct CachedEngine {
  std::shared_ptr<Tensor> gamma, eps_i, eps_a;
  bool complex_path = false;
  sisi_engine *e = nullptr;
};
std::vector<CachedEngine> &engines() {
  static std::vector<CachedEngine> v;
  return v;
}
}  // namespace

sisi_engine *engine_for(const std::shared_ptr<Tensor> &gamma, const std::shared_ptr<Tensor> &eps_i,
                        const std::shared_ptr<Tensor> &eps_a, bool complex_path) {
  for (auto &c : engines())
    if (c.gamma == gamma && c.eps_i == eps_i && c.eps_a == eps_a && c.complex_path == complex_path)
      return c.e;
  if (gamma->lens.size() != 3 || gamma->lens[1] != gamma->lens[2])
    throw StepError("CoulombVertex must be a [NG, Np, Np] tensor");
  if (eps_i->lens.size() != 1 || eps_a->lens.size() != 1)
    throw StepError("eigenenergies must be vectors");
  sisi_dims d{int(eps_i->lens[0]), int(eps_a->lens[0]), int(gamma->lens[0])};
  if (d.No + d.Nv != gamma->lens[1])
    throw StepError("CoulombVertex orbital length differs from No + Nv");
  if (eps_i->complex || eps_a->complex) throw StepError("eigenenergies must be real");
  const size_t n = size_t(gamma->size());
  std::vector<double> host;  // the vertex in the form the chosen path takes
  if (complex_path && !gamma->complex) {
    host.assign(2 * n, 0.0);
    for (size_t k = 0; k < n; ++k) host[2 * k] = gamma->values[k];
  } else if (!complex_path && gamma->complex) {
    host.resize(n);
    double imag_max = 0.0;
    for (size_t k = 0; k < n; ++k) {
      host[k] = gamma->values[2 * k];
      imag_max = std::max(imag_max, std::fabs(gamma->values[2 * k + 1]));
    }
    if (imag_max != 0.0)
      throw StepError("the CoulombVertex has an imaginary part: the real path needs a real "
                      "vertex (use complex: 1 in CoulombIntegralsFromVertex)");
  }
  const double *g = host.empty() ? gamma->values.data() : host.data();
  sisi_engine *e = nullptr;
  check(sisi_engine_create(&d, &e), "sisi_engine_create");
  if (complex_path) check(sisi_engine_set_vertex_complex(e, g), "sisi_engine_set_vertex_complex");
  else check(sisi_engine_set_vertex(e, g), "sisi_engine_set_vertex");
  check(sisi_engine_set_eigenenergies(e, eps_i->values.data(), eps_a->values.data()),
        "sisi_engine_set_eigenenergies");
  check(sisi_integrals_from_vertex(e), "sisi_integrals_from_vertex");
  engines().push_back({gamma, eps_i, eps_a, complex_path, e});
  return e;
}

void release_engines() {
  for (auto &c : engines()) sisi_engine_destroy(c.e);
  engines().clear();
}

}  // namespace sisi_host
