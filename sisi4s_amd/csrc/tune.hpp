// Kernel-variant autotuner of the GEMM families.
//
// A product is classified by its family (kernel family and pass count) and
// the power-of-two buckets of M, N and K.  The first product of a class
// that meets tuning enabled (SISI_TUNE=1, or sisi_tune_enable) times every
// candidate variant of its family on the device, on the real operands and a
// scratch output, and records the fastest; later products of the class use
// the recorded variant without timing.  Records are kept per process and
// persisted in a text cache: the file named by SISI_TUNE_CACHE, else
// tuning/gfx950.txt next to the engine library (shipped with the package),
// one line per class:
//
//   <family> <mb> <nb> <kb> <variant> <xcd> <us> <name>
//
// mb/nb/kb: ceil(log2) of M, N, K; xcd: 1 = XCD-contiguous tile ranges;
// us: measured microseconds of the winner (informational).
#pragma once

#include <string>
#include <vector>

namespace sisi {

// family ids: 3 / 1 split-BF16x3 / bf16 products; 103 / 101 the same with
// bf16 plane output; 64 / 32 FP64 / FP32 segmented products
enum class TuneFamily { B3 = 3, B1 = 1, B3P = 103, B1P = 101, Seg64 = 64, Seg32 = 32 };

struct TuneChoice {
  int variant = -1;
  int xcd = 1;
  double us = 0.0;
};

struct TuneKey {
  int family = 0, mb = 0, nb = 0, kb = 0;
  bool operator<(const TuneKey &o) const;
};

TuneKey tune_key(TuneFamily f, long long M, long long N, long long K);

// Recorded choice of a class, or variant -1.
TuneChoice tune_lookup(const TuneKey &k);
// Record (and append to the cache file when persistence is on).
void tune_record(const TuneKey &k, const TuneChoice &c, const std::string &name);
// Timing of unknown classes enabled for this process.
bool tune_enabled();
void tune_set_enabled(bool on);
// Path of the cache file in use ("" when none could be determined).
std::string tune_cache_path();
// Every record of this process (family mb nb kb variant xcd us name).
std::vector<std::string> tune_dump();

}  // namespace sisi
