// Host-side mirror of sisi4s' algorithm plugin surface for the CCSD / (T)
// path: a named-data map (src/Data.hpp:34-188), algorithms constructed by
// name from the input's steps (src/algorithms/Algorithm.hpp:104-183) with
// the reference's argument getters and defaults (Algorithm.cxx:37-140,
// 190-201, 289-295, 364-367).  The algorithms call the C ABI of
// include/sisi_engine.h; nothing here computes on the host.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sisi_engine.h"
#include "yaml_lite.hpp"

namespace sisi_host {

// Dense tensor, first index fastest (CTF global order); a complex tensor
// (the reference's Tensor<complex>) holds interleaved (re, im) pairs.
struct Tensor {
  std::vector<int64_t> lens;
  std::vector<double> values;
  bool complex = false;
  int64_t size() const {
    int64_t n = 1;
    for (auto l : lens) n *= l;
    return n;
  }
};

struct Datum {
  enum Type { None, TensorT, Real, Integer, Text } type = None;
  std::shared_ptr<Tensor> tensor;
  double real = 0;
  int64_t integer = 0;
  std::string text;
};

// Global named data ("$Name" symbols), as the reference's Data::dataMap.
class DataMap {
 public:
  Datum *find(const std::string &name) {
    auto it = data_.find(name);
    return it == data_.end() ? nullptr : &it->second;
  }
  void set(const std::string &name, Datum d) { data_[name] = std::move(d); }

 private:
  std::map<std::string, Datum> data_;
};

// Failure of a step (the reference throws DetailedException*).
struct StepError : std::runtime_error {
  explicit StepError(const std::string &m) : std::runtime_error(m) {}
};

// One argument of a step: a literal value or a $symbol.
struct Argument {
  std::string value;
  bool symbol = false;
};

// Records emitted to the output YAML (key path -> value text).
struct Emitted {
  std::vector<std::pair<std::string, std::string>> scalars;
  std::vector<std::vector<std::pair<std::string, std::string>>> iterations;
};

class Algorithm {
 public:
  Algorithm(std::map<std::string, Argument> in, std::map<std::string, Argument> out,
            DataMap &data)
      : in_(std::move(in)), out_(std::move(out)), data_(data) {}
  virtual ~Algorithm() = default;
  virtual std::string getName() = 0;
  virtual void run() = 0;

  bool isArgumentGiven(const std::string &n) const { return in_.count(n) || out_.count(n); }

  const Tensor &getTensorArgument(const std::string &n) { return *getTensorPtr(n); }
  std::shared_ptr<Tensor> getTensorPtr(const std::string &n);
  double getRealArgument(const std::string &n);
  double getRealArgument(const std::string &n, double def);
  int64_t getIntegerArgument(const std::string &n);
  int64_t getIntegerArgument(const std::string &n, int64_t def);
  std::string getTextArgument(const std::string &n);
  std::string getTextArgument(const std::string &n, const std::string &def);

  void setRealArgument(const std::string &n, double v);
  void allocatedTensorArgument(const std::string &n, std::shared_ptr<Tensor> t);

  Emitted emitted;

 protected:
  const Argument &input(const std::string &n) const;
  std::map<std::string, Argument> in_, out_;
  DataMap &data_;
};

using Factory = std::function<std::unique_ptr<Algorithm>(
    std::map<std::string, Argument>, std::map<std::string, Argument>, DataMap &)>;

std::map<std::string, Factory> &registry();

template <typename T>
struct Registrar {
  explicit Registrar(const std::string &name) {
    registry()[name] = [](std::map<std::string, Argument> in, std::map<std::string, Argument> out,
                          DataMap &d) { return std::unique_ptr<Algorithm>(new T(in, out, d)); };
  }
};

// Throws StepError with the engine's message on a non-zero status.
void check(int rc, const char *what);

// Engine shared by the steps that read the same vertex (uploaded once), on
// the real path or, with complex_path, on the complex one
// (sisi_engine_set_vertex_complex).  The real path needs a vertex without
// imaginary part.
sisi_engine *engine_for(const std::shared_ptr<Tensor> &gamma, const std::shared_ptr<Tensor> &eps_i,
                        const std::shared_ptr<Tensor> &eps_a, bool complex_path = false);
void release_engines();

}  // namespace sisi_host
