// Tuning cache of the GEMM kernel variants (see tune.hpp).
#include "tune.hpp"

#include <dlfcn.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

namespace sisi {

bool TuneKey::operator<(const TuneKey &o) const {
  if (family != o.family) return family < o.family;
  if (mb != o.mb) return mb < o.mb;
  if (nb != o.nb) return nb < o.nb;
  return kb < o.kb;
}

static int clog2(long long x) {
  int b = 0;
  while ((1LL << b) < x && b < 62) ++b;
  return b;
}

TuneKey tune_key(TuneFamily f, long long M, long long N, long long K) {
  TuneKey k;
  k.family = int(f);
  k.mb = clog2(M);
  k.nb = clog2(N);
  k.kb = clog2(K);
  return k;
}

namespace {

struct Store {
  std::mutex mu;
  std::map<TuneKey, TuneChoice> table;
  std::map<TuneKey, std::string> names;
  std::string path;
  bool loaded = false;
};

Store &store() {
  static Store s;
  return s;
}

std::string library_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void *>(&library_dir), &info) == 0 || info.dli_fname == nullptr)
    return "";
  std::string p(info.dli_fname);
  const size_t slash = p.rfind('/');
  return slash == std::string::npos ? std::string(".") : p.substr(0, slash);
}

// Load the cache once (caller holds the mutex).
void load_locked(Store &s) {
  if (s.loaded) return;
  s.loaded = true;
  const char *env = getenv("SISI_TUNE_CACHE");
  if (env != nullptr && *env) s.path = env;
  else {
    const std::string dir = library_dir();
    if (!dir.empty()) s.path = dir + "/tuning/gfx950.txt";
  }
  if (s.path.empty()) return;
  std::ifstream in(s.path);
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    TuneKey k;
    TuneChoice c;
    std::string name;
    if (ls >> k.family >> k.mb >> k.nb >> k.kb >> c.variant >> c.xcd >> c.us) {
      ls >> name;
      s.table[k] = c;  // later lines override earlier ones
      s.names[k] = name;
    }
  }
}

std::atomic<int> &enabled_flag() {
  static std::atomic<int> on{[] {
    const char *v = getenv("SISI_TUNE");
    return v != nullptr && *v && std::strcmp(v, "0") != 0 ? 1 : 0;
  }()};
  return on;
}

}  // namespace

TuneChoice tune_lookup(const TuneKey &k) {
  Store &s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  auto it = s.table.find(k);
  return it == s.table.end() ? TuneChoice{} : it->second;
}

void tune_record(const TuneKey &k, const TuneChoice &c, const std::string &name) {
  Store &s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  s.table[k] = c;
  s.names[k] = name;
  if (s.path.empty()) return;
  std::ofstream out(s.path, std::ios::app);
  if (!out) return;  // read-only installation: the record lives in this process only
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%d %d %d %d %d %d %.1f %s\n", k.family, k.mb, k.nb, k.kb,
                c.variant, c.xcd, c.us, name.c_str());
  out << buf;
}

bool tune_enabled() { return enabled_flag().load() != 0; }
void tune_set_enabled(bool on) { enabled_flag().store(on ? 1 : 0); }

std::string tune_cache_path() {
  Store &s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  return s.path;
}

std::vector<std::string> tune_dump() {
  Store &s = store();
  std::lock_guard<std::mutex> lock(s.mu);
  load_locked(s);
  std::vector<std::string> out;
  for (const auto &kv : s.table) {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%d %d %d %d %d %d %.1f %s", kv.first.family, kv.first.mb,
                  kv.first.nb, kv.first.kb, kv.second.variant, kv.second.xcd, kv.second.us,
                  s.names[kv.first].c_str());
    out.emplace_back(buf);
  }
  return out;
}

}  // namespace sisi
