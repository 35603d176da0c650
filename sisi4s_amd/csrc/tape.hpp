// Deferred launches for the lockstep batched solver.
//
// When an engine's `tape` is set, the launch helpers the CCSD residual and
// driver use (GEMM, strided copies, fills, axpby, the T1 T1 product, the
// amplitude estimate, the deterministic dot and energy reductions) append a
// record of the launch (kernel kind, template variant, eager grid, argument
// block) instead of launching.  The batched solver records one such tape
// per system and launches the i-th records of all systems together: one
// grouped launch per (step, kernel kind), each workgroup finding its system
// from a prefix sum of the systems' block counts.  The number of launches
// per iteration therefore does not grow with the number of systems.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"

#include <hip/hip_runtime.h>

namespace sisi {

struct CopyArgs {
  double *dst;
  const double *src;
  int nd;
  int64_t len[kMaxOrder];
  int64_t ds[kMaxOrder];
  int64_t ss[kMaxOrder];
  int64_t total;
  double alpha, beta;
  // tiled path: tile dims (0 = dst unit-stride dim, q = src unit-stride dim)
  int q;
  int64_t rest;  // product of the other dims
  unsigned gx;   // tiled path: tile blocks along x of the eager grid
};

// Arguments of the element-wise and reduction kernels (meaning per kind).
struct ElemArgs {
  double *y = nullptr;
  const double *x0 = nullptr, *x1 = nullptr, *x2 = nullptr, *x3 = nullptr;
  double a = 0.0, b = 0.0;
  int64_t n = 0, nelem = 0;
  int i0 = 0, i1 = 0, i2 = 0;
};

enum class OpKind : int {
  Gemm = 0,      // GemmArgs, sub = precision * 4 + AK * 2 + BNF; grid (tiles, ksplit, batch)
  CopyTiled,     // CopyArgs, block (32, 8)
  CopyGeneric,   // CopyArgs
  Fill,          // y[0..n) = a
  Axpby,         // y = a x0 + b y over n
  AddT1T1,       // y = x0 + a x1 (x) x2, i0 = Nv, i1 = No
  Estimate,      // R = y, T = x0, dT = x1, eps_i = x2, eps_a = x3, i0 = order, i1 = Nv,
                 // i2 = No, a = shift, n, nelem
  DotPartial,    // y[block] = partial sums of x0 . x1 over n
  SumPartials,   // y[c] = sum_b x0[b * i1 + c], c < i2, b < i0
  EnergyPartial, // y[2 block + {0,1}] = direct / exchange partials: T1 = x0, T2 = x1,
                 // Vabij = x2, Vbaij = x3, i0 = Nv, i1 = No
  GemmReduce,    // GemmArgs: split-K reduction of the preceding Gemm record, gx per
                 // batch entry
};

struct TapeOp {
  OpKind kind = OpKind::Fill;
  int sub = 0;
  unsigned gx = 1, gy = 1, gz = 1;  // eager grid of the launch
  GemmArgs g;
  CopyArgs c{};
  ElemArgs e;
  unsigned blocks() const { return gx * gy * gz; }
};

struct Tape {
  std::vector<TapeOp> ops;
  // argument records resident on the device (upload()): record i at
  // dev + off[i]; a tape without them has its records packed by each plan
  // that runs it
  DBuf dev;
  std::vector<size_t> off;
  void clear() {
    ops.clear();
    dev.reset();
    off.clear();
  }
  // Copy the argument records of every op to device memory (stream-ordered
  // on s, host-synchronous).
  void upload(Pool &pool, hipStream_t s);
};

// Partial-sum blocks of a taped reduction (fewer than the eager 1024: the
// grouped launch fills the device with many systems at once).
constexpr int kTapeRedBlocks = 64;

// Grouped launch of `n` records of one kind and variant: args[p] points to
// the argument block of record p (device memory), start[p] is its first
// block in the launch (start[n] = total blocks).
void launch_gemm_grouped(int sub, const GemmArgs *const *args, const unsigned *start, int n,
                         unsigned total, hipStream_t s);
// gxs[p]: the record's workgroups per batch entry
void launch_gemm_reduce_grouped(const GemmArgs *const *args, const unsigned *start,
                                const unsigned *gxs, int n, unsigned total, hipStream_t s);
void launch_copy_grouped(OpKind kind, const CopyArgs *const *args, const unsigned *start, int n,
                         unsigned total, hipStream_t s);
void launch_elem_grouped(OpKind kind, const ElemArgs *const *args, const unsigned *start, int n,
                         unsigned total, hipStream_t s);

// Lockstep schedule of several tapes (one per system): the i-th launches of
// the tapes are merged greedily into grouped launches of one (kind, variant)
// each, so that every tape's records run in their recorded order and the
// number of launches depends on the tapes' length, not on their number.
class GroupedPlan {
 public:
  GroupedPlan() = default;
  GroupedPlan(const GroupedPlan &) = delete;
  GroupedPlan &operator=(const GroupedPlan &) = delete;
  ~GroupedPlan() { reset(); }
  void build(Pool &pool, const std::vector<const Tape *> &tapes, hipStream_t s);
  // Run on s; with `graph`, the plan is captured in a hipGraph on its third
  // run and replayed from then on.
  void run(hipStream_t s, bool graph);
  void reset();
  size_t launches() const { return groups_.size(); }
  size_t records() const { return records_; }
  // Diagnostics: with a profile attached, run() launches eagerly with an
  // event pair around each grouped launch and adds the device time of each
  // (kind, variant) to profile[key] (key = kind * 16 + variant; milliseconds,
  // launches, records, blocks).
  struct ProfileEntry {
    double ms = 0.0, launches = 0.0, records = 0.0, blocks = 0.0;
  };
  std::vector<ProfileEntry> *profile = nullptr;
  // and per label (kind, variant and the shape of the group's largest record)
  std::map<std::string, ProfileEntry> *profile_shapes = nullptr;

 private:
  struct Group {
    OpKind kind;
    int sub;
    size_t ptrs, start, gxs;  // byte offsets in dev_ (gxs: GemmReduce only)
    int n;
    unsigned total;
    std::string label;  // diagnostics: shape of the largest record
  };
  std::vector<Group> groups_;
  DBuf dev_;
  hipGraphExec_t exec_ = nullptr;
  int runs_ = 0;
  size_t records_ = 0;
};

}  // namespace sisi
