// Contraction planner: maps one binary index-string product onto the MFMA
// GEMM family (see contract.hpp).
#include "contract.hpp"

#include <algorithm>
#include <map>
#include <vector>

#include "tensor_ops.hpp"

namespace sisi {

namespace {

struct Operand {
  const View *v;
  std::string lab;
  std::map<char, int64_t> stride;
};

Operand make_operand(const View &v, const std::string &lab, std::map<char, int64_t> &len) {
  SISI_REQUIRE(int(lab.size()) == v.order, SISI_ERR_INVALID,
               "contract: label string '" + lab + "' does not match tensor order");
  Operand o;
  o.v = &v;
  for (int k = 0; k < v.order; ++k) {
    const char c = lab[k];
    SISI_REQUIRE(o.stride.count(c) == 0, SISI_ERR_INVALID,
                 "contract: repeated label in '" + lab + "'");
    auto it = len.find(c);
    if (it == len.end()) len[c] = v.len[k];
    else
      SISI_REQUIRE(it->second == v.len[k], SISI_ERR_INVALID,
                   std::string("contract: length mismatch for label ") + c);
    if (v.len[k] == 1) continue;  // unit-length labels drop out
    o.lab.push_back(c);
    o.stride[c] = v.stride[k];
  }
  return o;
}

// labels of `group` ordered as they appear in `o`
std::string in_order(const Operand &o, const std::string &group) {
  std::string r;
  for (char c : o.lab)
    if (group.find(c) != std::string::npos) r.push_back(c);
  return r;
}

int64_t prod(const std::string &g, const std::map<char, int64_t> &len) {
  int64_t p = 1;
  for (char c : g) p *= len.at(c);
  return p;
}

// Is the ordered group one run (stride[g[i+1]] == stride[g[i]]*len[g[i]])?
bool fusable(const Operand &o, const std::string &g, const std::map<char, int64_t> &len,
             int64_t &s0) {
  if (g.empty()) {
    s0 = 0;
    return true;
  }
  s0 = o.stride.at(g[0]);
  for (size_t i = 0; i + 1 < g.size(); ++i)
    if (o.stride.at(g[i + 1]) != o.stride.at(g[i]) * len.at(g[i])) return false;
  return true;
}

// Split an ordered group into runs that are fusable in o.
std::vector<std::pair<int64_t, int64_t>> runs(const Operand &o, const std::string &g,
                                              const std::map<char, int64_t> &len) {
  std::vector<std::pair<int64_t, int64_t>> r;  // (length, stride)
  for (size_t i = 0; i < g.size(); ++i) {
    const int64_t l = len.at(g[i]), s = o.stride.at(g[i]);
    if (!r.empty() && r.back().second * r.back().first == s) r.back().first *= l;
    else r.push_back({l, s});
  }
  return r;
}

// Operand usable as a GEMM operand with groups (g1, g2, batch)?
bool operand_ok(const Operand &o, const std::string &g1, const std::string &g2,
                const std::string &gb, const std::map<char, int64_t> &len) {
  int64_t s1, s2, sb;
  if (!fusable(o, g1, len, s1) || !fusable(o, g2, len, s2) || !fusable(o, gb, len, sb))
    return false;
  const bool u1 = g1.empty() || s1 == 1;
  const bool u2 = g2.empty() || s2 == 1;
  return u1 || u2;
}

// Dense copy of an operand in label order `order`.
View densify(Engine &e, const Operand &o, const std::string &order,
             const std::map<char, int64_t> &len, DBuf &store) {
  int64_t n = 1;
  for (char c : order) n *= len.at(c);
  // Constant operands keep their permuted copies for the whole residual.
  std::string key;
  const bool cached = e.is_frozen(o.v->p);
  if (cached) {
    key = std::to_string(reinterpret_cast<uintptr_t>(o.v->p));
    for (char c : order) key += ":" + std::to_string(len.at(c)) + "/" + std::to_string(o.stride.at(c));
  }
  View d;
  d.order = int(order.size());
  if (cached) {
    auto it = e.perm_cache.find(key);
    if (it != e.perm_cache.end()) {
      d.p = it->second.get();
      int64_t st = 1;
      for (int k = 0; k < d.order; ++k) {
        d.len[k] = len.at(order[k]);
        d.stride[k] = st;
        st *= d.len[k];
      }
      return d;
    }
  }
  store = DBuf(&e.pool, size_t(std::max<int64_t>(n, 1)));
  d.p = store.get();
  View s;
  s.p = o.v->p;
  s.order = d.order;
  int64_t st = 1;
  for (int k = 0; k < d.order; ++k) {
    const char c = order[k];
    d.len[k] = s.len[k] = len.at(c);
    d.stride[k] = st;
    s.stride[k] = o.stride.at(c);
    st *= len.at(c);
  }
  if (d.order == 0) {
    d.order = s.order = 1;
    d.len[0] = s.len[0] = 1;
    d.stride[0] = s.stride[0] = 1;
  }
  copy(e, d, s);
  if (cached) e.perm_cache[key] = std::move(store);
  return d;
}

Operand operand_from_dense(const View &d, const std::string &order) {
  Operand o;
  o.v = &d;
  o.lab = order;
  for (size_t k = 0; k < order.size(); ++k) o.stride[order[k]] = d.stride[k];
  return o;
}

}  // namespace

void contract(Engine &e, double alpha, const View &Av, const std::string &ia, const View &Bv,
              const std::string &ib, double beta, const View &Cv, const std::string &ic,
              int precision) {
  if (precision < 0) precision = e.precision;
  std::map<char, int64_t> len;
  Operand A = make_operand(Av, ia, len);
  Operand B = make_operand(Bv, ib, len);
  Operand C = make_operand(Cv, ic, len);
  if (Cv.size() == 0) return;

  std::string gM, gN, gK, gB;
  auto has = [](const Operand &o, char c) { return o.stride.count(c) > 0; };
  for (auto &kv : len) {
    const char c = kv.first;
    if (kv.second == 1) continue;
    const bool a = has(A, c), b = has(B, c), cc = has(C, c);
    if (a && b && cc) gB.push_back(c);
    else if (a && cc) gM.push_back(c);
    else if (b && cc) gN.push_back(c);
    else if (a && b) gK.push_back(c);
    else
      throw Error(SISI_ERR_INVALID, std::string("contract: label ") + c +
                                        " must appear in two of the three operands (" + ia +
                                        "," + ib + "->" + ic + ")");
  }
  std::string Mord = in_order(A, gM), Nord = in_order(B, gN), Bord = in_order(C, gB);
  std::string KA = in_order(A, gK), KB = in_order(B, gK), Kord;
  // The epilogue's lanes run along the first M (or, swapped, N) label.  When
  // C's unit-stride label is an M or N label that does not lead its group,
  // the group follows C's strides instead, provided the operand that then
  // needs a permuted copy is small against C (e.g. "jika,bk->abij": the
  // output is Nv/No times larger than V_ijka, and led by j its writes would
  // be Nv^2 No elements apart).
  {
    auto lead_unit = [&](const std::string &grp) {
      return !grp.empty() && C.stride.at(grp[0]) == 1;
    };
    auto by_c = [&](std::string grp) {
      std::stable_sort(grp.begin(), grp.end(),
                       [&](char x, char y) { return C.stride.at(x) < C.stride.at(y); });
      return grp;
    };
    const int64_t nA = prod(A.lab, len), nB = prod(B.lab, len), nC = prod(C.lab, len);
    if (!lead_unit(Mord) && !lead_unit(Nord)) {
      const std::string cm = by_c(Mord), cn = by_c(Nord);
      if (lead_unit(cm) && 4 * nA <= nC) Mord = cm;
      else if (lead_unit(cn) && 4 * nB <= nC) Nord = cn;
    }
  }

  // operands: choose the k order that needs the fewest copies
  DBuf storeA, storeB, storeC;
  View dA, dB;
  const bool a1 = operand_ok(A, Mord, KA, Bord, len), b1 = operand_ok(B, KA, Nord, Bord, len);
  const bool a2 = operand_ok(A, Mord, KB, Bord, len), b2 = operand_ok(B, KB, Nord, Bord, len);
  bool permA = false, permB = false;
  if (a1 && b1) Kord = KA;
  else if (a2 && b2) Kord = KB;
  else if (a1) {
    Kord = KA;
    permB = true;
  } else if (b2) {
    Kord = KB;
    permA = true;
  } else {
    Kord = KA;
    permA = permB = true;
  }
  if (permA) {
    dA = densify(e, A, Mord + Kord + Bord, len, storeA);
    A = operand_from_dense(dA, Mord + Kord + Bord);
  }
  if (permB) {
    dB = densify(e, B, Kord + Nord + Bord, len, storeB);
    B = operand_from_dense(dB, Kord + Nord + Bord);
  }

  // output: direct when each group is <= 3 runs and the batch group is one run
  auto rM = runs(C, Mord, len), rN = runs(C, Nord, len);
  int64_t sCb = 0;
  bool c_direct = rM.size() <= 3 && rN.size() <= 3 && fusable(C, Bord, len, sCb);
  View tmpC;
  Operand Cw = C;
  if (!c_direct) {
    const std::string order = Mord + Nord + Bord;
    int64_t n = prod(order, len);
    storeC = DBuf(&e.pool, size_t(n));
    tmpC.p = storeC.get();
    tmpC.order = int(order.size());
    int64_t st = 1;
    for (int k = 0; k < tmpC.order; ++k) {
      tmpC.len[k] = len[order[k]];
      tmpC.stride[k] = st;
      st *= tmpC.len[k];
    }
    Cw = operand_from_dense(tmpC, order);
    rM = runs(Cw, Mord, len);
    rN = runs(Cw, Nord, len);
    fusable(Cw, Bord, len, sCb);
  }

  // lanes of the GEMM epilogue run along m: put C's unit-stride label there
  bool swap = false;
  if (!rN.empty() && rN[0].second == 1 && !(!rM.empty() && rM[0].second == 1)) swap = true;

  GemmArgs g;
  int64_t sAm, sAk, sBk, sBn, sAb, sBb;
  fusable(A, Mord, len, sAm);
  fusable(A, Kord, len, sAk);
  fusable(A, Bord, len, sAb);
  fusable(B, Kord, len, sBk);
  fusable(B, Nord, len, sBn);
  fusable(B, Bord, len, sBb);
  const int64_t M = prod(Mord, len), N = prod(Nord, len), K = prod(Kord, len),
                nb = prod(Bord, len);
  SISI_REQUIRE(M < (int64_t(1) << 31) && N < (int64_t(1) << 31) && K < (int64_t(1) << 31),
               SISI_ERR_INVALID, "contract: GEMM dimension too large");
  if (Mord.empty()) sAm = 0;
  if (Kord.empty()) sAk = sBk = 0;
  if (Nord.empty()) sBn = 0;
  auto set_out = [&](const std::vector<std::pair<int64_t, int64_t>> &rm,
                     const std::vector<std::pair<int64_t, int64_t>> &rn) {
    int64_t l[3] = {1, 1, 1}, s[3] = {0, 0, 0};
    for (size_t k = 0; k < rm.size(); ++k) {
      l[k] = rm[k].first;
      s[k] = rm[k].second;
    }
    g.dm0 = int(std::max<int64_t>(l[0], 1));
    g.dm1 = int(std::max<int64_t>(l[1], 1));
    g.cm0 = s[0];
    g.cm1 = s[1];
    g.cm2 = s[2];
    int64_t ln[3] = {1, 1, 1}, sn[3] = {0, 0, 0};
    for (size_t k = 0; k < rn.size(); ++k) {
      ln[k] = rn[k].first;
      sn[k] = rn[k].second;
    }
    g.dn0 = int(std::max<int64_t>(ln[0], 1));
    g.dn1 = int(std::max<int64_t>(ln[1], 1));
    g.cn0 = sn[0];
    g.cn1 = sn[1];
    g.cn2 = sn[2];
  };
  if (!swap) {
    g.A = A.v->p;
    g.B = B.v->p;
    g.M = int(M);
    g.N = int(N);
    g.sam = sAm;
    g.sak = sAk;
    g.sbk = sBk;
    g.sbn = sBn;
    g.bsa = sAb;
    g.bsb = sBb;
    set_out(rM, rN);
  } else {
    g.A = B.v->p;
    g.B = A.v->p;
    g.M = int(N);
    g.N = int(M);
    g.sam = sBn;
    g.sak = sBk;
    g.sbk = sAk;
    g.sbn = sAm;
    g.bsa = sBb;
    g.bsb = sAb;
    set_out(rN, rM);
  }
  g.K = int(K);
  g.batch = int(nb);
  g.bsc = sCb;
  g.C = Cw.v->p;
  g.alpha = alpha;
  g.beta = c_direct ? beta : 0.0;
  gemm(e, precision, g);

  if (!c_direct) {
    // C[ic] = tmp[Mord Nord Bord] + beta * C[ic]
    View src, dst;
    src.p = tmpC.p;
    dst.p = Cv.p;
    const std::string order = Mord + Nord + Bord;
    src.order = dst.order = int(order.size());
    for (int k = 0; k < src.order; ++k) {
      src.len[k] = dst.len[k] = tmpC.len[k];
      src.stride[k] = tmpC.stride[k];
      dst.stride[k] = C.stride.at(order[k]);
    }
    copy(e, dst, src, 1.0, beta);
  }
}

void ctr(Engine &e, double alpha, const char *spec, const View &A, const View &B, double beta,
         const View &C, int precision) {
  std::string s(spec);
  const size_t comma = s.find(','), arrow = s.find("->");
  SISI_REQUIRE(comma != std::string::npos && arrow != std::string::npos && comma < arrow,
               SISI_ERR_INVALID, "ctr: bad spec " + s);
  contract(e, alpha, A, s.substr(0, comma), B, s.substr(comma + 1, arrow - comma - 1), beta,
           C, s.substr(arrow + 2), precision);
}

void sum(Engine &e, double alpha, const View &A, const std::string &ia, double beta,
         const View &C, const std::string &ic) {
  SISI_REQUIRE(int(ia.size()) == A.order && int(ic.size()) == C.order && ia.size() == ic.size(),
               SISI_ERR_INVALID, "sum: label mismatch");
  View src;
  src.p = A.p;
  src.order = C.order;
  for (int k = 0; k < C.order; ++k) {
    const size_t pos = ia.find(ic[k]);
    SISI_REQUIRE(pos != std::string::npos, SISI_ERR_INVALID, "sum: label not found");
    SISI_REQUIRE(A.len[pos] == C.len[k], SISI_ERR_INVALID, "sum: length mismatch");
    src.len[k] = C.len[k];
    src.stride[k] = A.stride[pos];
  }
  copy(e, C, src, alpha, beta);
}

}  // namespace sisi
