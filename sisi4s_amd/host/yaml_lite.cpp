#include "yaml_lite.hpp"

#include <fstream>
#include <sstream>

namespace sisi_host {

namespace {

struct Line {
  int indent;
  std::string text;  // without indentation and comments, right-trimmed
  int number;
};

std::string rtrim(const std::string &s) {
  size_t e = s.find_last_not_of(" \t\r");
  return e == std::string::npos ? "" : s.substr(0, e + 1);
}

std::string trim(const std::string &s) {
  size_t b = s.find_first_not_of(" \t\r");
  if (b == std::string::npos) return "";
  return rtrim(s.substr(b));
}

// strip a comment that starts with '#' outside quotes (and at a word start)
std::string strip_comment(const std::string &s) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t'))
      return s.substr(0, i);
  }
  return s;
}

class Parser {
 public:
  explicit Parser(const std::string &text) {
    std::istringstream in(text);
    std::string raw;
    int n = 0;
    while (std::getline(in, raw)) {
      ++n;
      std::string s = rtrim(strip_comment(raw));
      if (trim(s).empty()) continue;
      if (trim(s) == "---" || trim(s) == "...") continue;
      // a flow collection ({...} or [...]) may continue on the next lines
      const int first = n;
      while (flow_depth(s) > 0 && std::getline(in, raw)) {
        ++n;
        s += " " + trim(strip_comment(raw));
      }
      if (flow_depth(s) != 0)
        throw YamlError("line " + std::to_string(first) + ": unbalanced { } or [ ]");
      int ind = 0;
      while (ind < int(s.size()) && s[ind] == ' ') ++ind;
      if (ind < int(s.size()) && s[ind] == '\t')
        throw YamlError("line " + std::to_string(n) + ": tabs are not allowed for indentation");
      lines_.push_back({ind, s.substr(ind), first});
    }
  }

  YamlNode parse() {
    if (lines_.empty()) return YamlNode{};
    YamlNode r = block(lines_[0].indent);
    if (pos_ < lines_.size())
      throw YamlError("line " + std::to_string(lines_[pos_].number) + ": unexpected indentation");
    return r;
  }

 private:
  std::vector<Line> lines_;
  size_t pos_ = 0;
  std::map<std::string, YamlNode> anchors_;

  static bool is_seq_item(const std::string &t) { return t == "-" || t.rfind("- ", 0) == 0; }

  // open { and [ minus closed ones, outside quotes
  static int flow_depth(const std::string &t) {
    bool sq = false, dq = false;
    int d = 0;
    for (char c : t) {
      if (c == '\'' && !dq) sq = !sq;
      else if (c == '"' && !sq) dq = !dq;
      else if (!sq && !dq && (c == '{' || c == '[')) ++d;
      else if (!sq && !dq && (c == '}' || c == ']')) --d;
    }
    return d;
  }

  // Flow collections: [a, b], {k: v, k2: v2}, nested; plain items go through
  // scalar() (quotes, anchors, aliases).
  YamlNode flow(const std::string &s, size_t &i, int lineno) {
    auto skip = [&]() {
      while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
    };
    auto fail = [&](const std::string &m) -> YamlError {
      return YamlError("line " + std::to_string(lineno) + ": " + m);
    };
    skip();
    YamlNode n;
    if (i < s.size() && (s[i] == '[' || s[i] == '{')) {
      const bool is_map = s[i] == '{';
      const char close = is_map ? '}' : ']';
      n.type = is_map ? YamlNode::Map : YamlNode::Seq;
      ++i;
      skip();
      if (i < s.size() && s[i] == close) {
        ++i;
        return n;
      }
      while (true) {
        if (is_map) {
          // key up to ':' at depth 0
          size_t k0 = i;
          bool sq = false, dq = false;
          while (i < s.size() && !(s[i] == ':' && !sq && !dq)) {
            if (s[i] == '\'' && !dq) sq = !sq;
            else if (s[i] == '"' && !sq) dq = !dq;
            else if (!sq && !dq && (s[i] == ',' || s[i] == close)) throw fail("expected 'key: value' in {}");
            ++i;
          }
          if (i >= s.size()) throw fail("unterminated {");
          std::string key = trim(s.substr(k0, i - k0));
          if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'')) key = key.substr(1, key.size() - 2);
          ++i;  // ':'
          YamlNode v = flow(s, i, lineno);
          for (auto &kv : n.map)
            if (kv.first == key) throw fail("duplicate key " + key);
          n.map.push_back({key, v});
        } else {
          n.seq.push_back(flow(s, i, lineno));
        }
        skip();
        if (i >= s.size()) throw fail(std::string("unterminated ") + (is_map ? "{" : "["));
        if (s[i] == ',') {
          ++i;
          continue;
        }
        if (s[i] == close) {
          ++i;
          break;
        }
        throw fail(std::string("expected ',' or '") + close + "'");
      }
      return n;
    }
    // plain or quoted item up to ',' ']' '}' at depth 0
    size_t b = i;
    bool sq = false, dq = false;
    while (i < s.size()) {
      const char c = s[i];
      if (c == '\'' && !dq) sq = !sq;
      else if (c == '"' && !sq) dq = !dq;
      else if (!sq && !dq && (c == ',' || c == ']' || c == '}')) break;
      ++i;
    }
    return scalar(s.substr(b, i - b), lineno);
  }

  // position of the "key:" separator (colon followed by space or end), -1 if none
  static int key_colon(const std::string &t) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < t.size(); ++i) {
      const char c = t[i];
      if (c == '\'' && !dq) sq = !sq;
      else if (c == '"' && !sq) dq = !dq;
      else if (c == ':' && !sq && !dq && (i + 1 == t.size() || t[i + 1] == ' ')) return int(i);
    }
    return -1;
  }

  YamlNode scalar(const std::string &v, int lineno) {
    std::string s = trim(v);
    YamlNode n;
    if (s.empty()) return n;
    if (s[0] == '*') {
      auto it = anchors_.find(s.substr(1));
      if (it == anchors_.end())
        throw YamlError("line " + std::to_string(lineno) + ": unknown alias " + s);
      return it->second;
    }
    std::string anchor;
    if (s[0] == '&') {
      size_t sp = s.find(' ');
      anchor = s.substr(1, sp == std::string::npos ? std::string::npos : sp - 1);
      s = sp == std::string::npos ? "" : trim(s.substr(sp + 1));
    }
    if (!s.empty() && (s[0] == '[' || s[0] == '{')) {
      size_t i = 0;
      n = flow(s, i, lineno);
      if (!trim(s.substr(i)).empty())
        throw YamlError("line " + std::to_string(lineno) + ": text after a flow collection");
    } else if (!s.empty()) {
      n.type = YamlNode::Scalar;
      if ((s[0] == '"' || s[0] == '\'') && s.size() >= 2 && s.back() == s[0]) {
        n.scalar = s.substr(1, s.size() - 2);
        n.quoted = true;
      } else {
        n.scalar = s;
      }
    }
    if (!anchor.empty()) anchors_[anchor] = n;
    return n;
  }

  YamlNode block(int indent) {
    if (pos_ >= lines_.size()) return YamlNode{};
    if (is_seq_item(lines_[pos_].text)) return seq(indent);
    return map(indent);
  }

  YamlNode seq(int indent) {
    YamlNode n;
    n.type = YamlNode::Seq;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent && is_seq_item(lines_[pos_].text)) {
      Line &l = lines_[pos_];
      const std::string rest = l.text.substr(1);  // after '-'
      const std::string r = trim(rest);
      if (r.empty()) {
        ++pos_;
        if (pos_ < lines_.size() && lines_[pos_].indent > indent)
          n.seq.push_back(block(lines_[pos_].indent));
        else
          n.seq.push_back(YamlNode{});
      } else if (key_colon(r) >= 0) {
        // "- key: value": the item is a map whose keys sit at the column of "key"
        const int inner = indent + 1 + int(rest.find_first_not_of(' '));
        l.indent = inner;
        l.text = r;
        n.seq.push_back(map(inner));
      } else {
        n.seq.push_back(scalar(r, l.number));
        ++pos_;
      }
    }
    return n;
  }

  YamlNode map(int indent) {
    YamlNode n;
    n.type = YamlNode::Map;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent && !is_seq_item(lines_[pos_].text)) {
      const Line &l = lines_[pos_];
      const int c = key_colon(l.text);
      if (c < 0) throw YamlError("line " + std::to_string(l.number) + ": expected 'key: value'");
      std::string key = trim(l.text.substr(0, c));
      if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'')) key = key.substr(1, key.size() - 2);
      std::string val = trim(l.text.substr(c + 1));
      const int lineno = l.number;
      ++pos_;
      YamlNode child;
      std::string anchor;
      if (!val.empty() && val[0] == '&' && val.find(' ') == std::string::npos) {
        anchor = val.substr(1);  // anchor on a nested block
        val.clear();
      }
      if (val.empty()) {
        if (pos_ < lines_.size() && lines_[pos_].indent > indent) child = block(lines_[pos_].indent);
        else if (pos_ < lines_.size() && lines_[pos_].indent == indent &&
                 is_seq_item(lines_[pos_].text))
          child = seq(indent);  // "key:\n- item" at the same indentation
      } else {
        child = scalar(val, lineno);
      }
      if (!anchor.empty()) anchors_[anchor] = child;
      for (auto &kv : n.map)
        if (kv.first == key) throw YamlError("line " + std::to_string(lineno) + ": duplicate key " + key);
      n.map.push_back({key, child});
    }
    return n;
  }
};

}  // namespace

YamlNode parse_yaml(const std::string &text) { return Parser(text).parse(); }

YamlNode parse_yaml_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw YamlError("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_yaml(ss.str());
}

}  // namespace sisi_host
