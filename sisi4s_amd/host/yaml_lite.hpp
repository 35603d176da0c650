// Minimal YAML reader for sisi4s input files and cc4s data headers: block
// maps and sequences, flow collections ([...], {...}, nested, possibly over
// several lines), scalars, $Symbol references and anchors/aliases (&x / *x),
// as in integration-tests/tests/cc4s/hf/in.yaml and
// integration-tests/tests/cc4s/ueg/rs1.0-7occ-26virt/in.yaml.  The reference parses the same
// files with yaml-cpp (src/Parser.cxx:21-72).
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sisi_host {

struct YamlNode {
  enum Type { Null, Scalar, Map, Seq } type = Null;
  std::string scalar;
  bool quoted = false;
  std::vector<std::pair<std::string, YamlNode>> map;
  std::vector<YamlNode> seq;

  const YamlNode *get(const std::string &key) const {
    for (auto &kv : map)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

struct YamlError : std::runtime_error {
  explicit YamlError(const std::string &m) : std::runtime_error(m) {}
};

YamlNode parse_yaml(const std::string &text);
YamlNode parse_yaml_file(const std::string &path);

}  // namespace sisi_host
