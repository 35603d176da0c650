// Step runner for sisi4s-style YAML inputs (src/Sisi4s.cxx:22-129): each step
// {name, in, out, fallible, disable} constructs the named algorithm, runs it,
// times it and records its results; `$Name` values are data symbols.  The
// output YAML keeps the reference's per-step keys (name, in, out, realtime,
// iterations[].energy).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "algorithm.hpp"
#include "yaml_lite.hpp"

using namespace sisi_host;

namespace {

struct Step {
  std::string name, note;
  std::map<std::string, Argument> in, out;
  bool fallible = false, disabled = false;
};

bool truthy(const YamlNode *n) {
  if (!n || n->type != YamlNode::Scalar) return false;
  const std::string &s = n->scalar;
  return s == "1" || s == "true" || s == "True" || s == "yes";
}

std::map<std::string, Argument> arguments(const YamlNode *n, const std::string &where) {
  std::map<std::string, Argument> r;
  if (!n || n->type == YamlNode::Null) return r;
  if (n->type != YamlNode::Map) throw YamlError(where + " must be a map");
  for (auto &kv : n->map) {
    if (kv.second.type != YamlNode::Scalar)
      throw YamlError(where + "." + kv.first + " must be a scalar or $symbol");
    Argument a;
    a.value = kv.second.scalar;
    if (!kv.second.quoted && !a.value.empty() && a.value[0] == '$') {
      a.symbol = true;
      a.value = a.value.substr(1);
    }
    r[kv.first] = a;
  }
  return r;
}

std::vector<Step> read_steps(const std::string &path) {
  YamlNode root = parse_yaml_file(path);
  if (root.type != YamlNode::Seq) throw YamlError("the input must be a sequence of steps");
  std::vector<Step> steps;
  for (auto &item : root.seq) {
    if (item.type != YamlNode::Map) throw YamlError("each step must be a map");
    const YamlNode *name = item.get("name");
    if (!name || name->type != YamlNode::Scalar) throw YamlError("step without a name");
    Step s;
    s.name = name->scalar;
    s.in = arguments(item.get("in"), s.name + ".in");
    s.out = arguments(item.get("out"), s.name + ".out");
    s.fallible = truthy(item.get("fallible"));
    s.disabled = truthy(item.get("disable")) || (item.get("enable") && !truthy(item.get("enable")));
    if (const YamlNode *nt = item.get("note")) s.note = nt->scalar;
    steps.push_back(s);
  }
  return steps;
}

std::string yaml_str(const std::string &s) {
  std::string r = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') r += '\\';
    if (c == '\n') {
      r += "\\n";
      continue;
    }
    r += c;
  }
  return r + "\"";
}

std::string flow(const std::map<std::string, Argument> &m) {
  std::string r = "{";
  bool first = true;
  for (auto &kv : m) {
    if (!first) r += ", ";
    first = false;
    r += kv.first + ": " + (kv.second.symbol ? "$" + kv.second.value : yaml_str(kv.second.value));
  }
  return r + "}";
}

}  // namespace

int main(int argc, char **argv) {
  std::string in, out = "sisi4s.out.yaml";
  bool dry = false, list = false;
  for (int a = 1; a < argc; ++a) {
    if ((!strcmp(argv[a], "-i") || !strcmp(argv[a], "--in")) && a + 1 < argc) in = argv[++a];
    else if ((!strcmp(argv[a], "-o") || !strcmp(argv[a], "--out")) && a + 1 < argc) out = argv[++a];
    else if (!strcmp(argv[a], "--dry")) dry = true;
    else if (!strcmp(argv[a], "--list")) list = true;
    else {
      std::cerr << "usage: sisi4s-mi355x -i input.yaml [-o output.yaml] [--dry] [--list]\n";
      return 2;
    }
  }
  if (list) {
    for (auto &kv : registry()) std::cout << kv.first << "\n";
    std::cout << "Nop\n";
    return 0;
  }
  if (in.empty()) {
    std::cerr << "no input file (-i)\n";
    return 2;
  }
  std::vector<Step> steps;
  try {
    steps = read_steps(in);
  } catch (const std::exception &ex) {
    std::cerr << "input error: " << ex.what() << "\n";
    return 1;
  }
  for (auto &s : steps)
    if (s.name != "Nop" && !registry().count(s.name)) {
      std::cerr << "Algorithm not implemented: " << s.name << "\n";
      return 1;
    }
  if (dry) {
    for (auto &s : steps)
      std::cout << s.name << " in=" << flow(s.in) << " out=" << flow(s.out)
                << (s.disabled ? " (disabled)" : "") << (s.fallible ? " (fallible)" : "") << "\n";
    return 0;
  }
  DataMap data;
  std::ostringstream y;
  y << "steps:\n";
  int rc = 0;
  for (auto &s : steps) {
    if (s.disabled || s.name == "Nop") continue;
    y << "  - name: " << s.name << "\n";
    if (!s.note.empty()) y << "    note: " << yaml_str(s.note) << "\n";
    y << "    in: " << flow(s.in) << "\n    out: " << flow(s.out) << "\n";
    std::cout << "step: " << s.name << std::endl;
    auto t0 = std::chrono::steady_clock::now();
    try {
      auto alg = registry()[s.name](s.in, s.out, data);
      alg->run();
      const double dt =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      y << "    realtime: " << dt << "\n";
      for (auto &kv : alg->emitted.scalars) y << "    " << kv.first << ": " << kv.second << "\n";
      if (!alg->emitted.iterations.empty()) {
        y << "    iterations:\n";
        for (auto &it : alg->emitted.iterations) {
          y << "      - {";
          for (size_t k = 0; k < it.size(); ++k)
            y << (k ? ", " : "") << it[k].first << ": "
              << (it[k].first == "energy" ? "{value: " + it[k].second + "}" : it[k].second);
          y << "}\n";
        }
      }
      for (auto &kv : s.out) {
        Datum *d = kv.second.symbol ? data.find(kv.second.value) : nullptr;
        if (d && d->type == Datum::Real) {
          std::ostringstream v;
          v.precision(17);
          v << d->real;
          std::cout << "  " << kv.first << " = " << v.str() << std::endl;
          y << "    " << kv.first << ": " << v.str() << "\n";
        }
      }
    } catch (const std::exception &ex) {
      y << "    error: " << yaml_str(ex.what()) << "\n";
      std::cerr << s.name << " failed: " << ex.what() << "\n";
      if (!s.fallible) {
        rc = 1;
        break;
      }
    }
  }
  release_engines();
  std::ofstream f(out);
  f << y.str();
  return rc;
}
