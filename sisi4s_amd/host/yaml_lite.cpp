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
      int ind = 0;
      while (ind < int(s.size()) && s[ind] == ' ') ++ind;
      if (ind < int(s.size()) && s[ind] == '\t')
        throw YamlError("line " + std::to_string(n) + ": tabs are not allowed for indentation");
      lines_.push_back({ind, s.substr(ind), n});
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
    if (!s.empty() && s[0] == '[') {
      if (s.back() != ']') throw YamlError("line " + std::to_string(lineno) + ": unterminated [");
      n.type = YamlNode::Seq;
      std::string body = s.substr(1, s.size() - 2);
      std::stringstream ss(body);
      std::string item;
      while (std::getline(ss, item, ','))
        if (!trim(item).empty()) n.seq.push_back(scalar(item, lineno));
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
