// CCSD hot path on the device.  Every contraction below names the reference
// line it evaluates; operands are FP64 in HBM and the contraction precision
// is the engine's (FP64, FP32, or BF16x3 for the ladder).
#include "ccsd.hpp"

#include <cmath>

#include "contract.hpp"
#include "tensor_ops.hpp"

namespace sisi {

namespace {
inline View V2(double *p, int64_t a, int64_t b) { return dense(p, {a, b}); }
inline View V3(double *p, int64_t a, int64_t b, int64_t c) { return dense(p, {a, b, c}); }
inline View V4(double *p, int64_t a, int64_t b, int64_t c, int64_t d) {
  return dense(p, {a, b, c, d});
}
inline View V1(double *p, int64_t a) { return dense(p, {a}); }
}  // namespace

void set_precision(Engine &e, int precision) {
  SISI_REQUIRE(precision == SISI_FP64 || precision == SISI_FP32 || precision == SISI_BF16X3,
               SISI_ERR_INVALID, "unknown precision");
  e.precision = precision == SISI_FP64 ? SISI_FP64 : SISI_FP32;
  e.prec_ladder = precision;
}

void integrals_from_vertex(Engine &e) {
  e.require_inputs();
  const int64_t o = e.No, v = e.Nv, G = e.NG, P = e.Np;
  // slices (CoulombIntegralsFromVertex.cxx:121-136; holes are p < No)
  View gam = V3(e.gamma.get(), G, P, P);
  e.Gij = DBuf(&e.pool, G * o * o);
  e.Gai = DBuf(&e.pool, G * v * o);
  e.Gab = DBuf(&e.pool, G * v * v);
  copy(e, V3(e.Gij.get(), G, o, o), subview(gam, {0, 0, 0}, {G, o, o}));
  copy(e, V3(e.Gai.get(), G, v, o), subview(gam, {0, o, 0}, {G, P, o}));
  copy(e, V3(e.Gab.get(), G, v, v), subview(gam, {0, o, o}, {G, P, P}));
  View Gij = V3(e.Gij.get(), G, o, o), Gai = V3(e.Gai.get(), G, v, o),
       Gab = V3(e.Gab.get(), G, v, v);
  e.Vaibj = DBuf(&e.pool, v * o * v * o);
  e.Vabij = DBuf(&e.pool, v * v * o * o);
  e.Vijkl = DBuf(&e.pool, o * o * o * o);
  e.Vijka = DBuf(&e.pool, o * o * o * v);
  ctr(e, 1.0, "Gab,Gij->aibj", Gab, Gij, 0.0, V4(e.Vaibj.get(), v, o, v, o), SISI_FP64);  // :395
  ctr(e, 1.0, "Gai,Gbj->abij", Gai, Gai, 0.0, V4(e.Vabij.get(), v, v, o, o), SISI_FP64);  // :402
  ctr(e, 1.0, "Gik,Gjl->ijkl", Gij, Gij, 0.0, V4(e.Vijkl.get(), o, o, o, o), SISI_FP64);  // :409
  ctr(e, 1.0, "Gik,Gaj->ijka", Gij, Gai, 0.0, V4(e.Vijka.get(), o, o, o, v), SISI_FP64);  // :416
  // constants of the residual and the energy
  View V = V4(e.Vabij.get(), v, v, o, o);
  e.Vbaij = DBuf(&e.pool, v * v * o * o);
  e.Uabij = DBuf(&e.pool, v * v * o * o);
  sum(e, 1.0, V, "baij", 0.0, V4(e.Vbaij.get(), v, v, o, o), "abij");
  // U = 2 V_abij - V_baij: the combination (2 V[cdkl] - V[dckl]) of
  // CcsdEnergyFromCoulombIntegrals.cxx:146-147, 172-173, 462-463
  axpby(e, v * v * o * o, 2.0, e.Vabij.get(), 0.0, e.Uabij.get());
  axpby(e, v * v * o * o, -1.0, e.Vbaij.get(), 1.0, e.Uabij.get());
  e.Vt.reset();
  e.Vt2.reset();
  e.VP.reset();
  e.have_integrals = true;
  e.sync();
}

void ccsd_residual(Engine &e, int it, bool initial_doubles, int slice_size, const double *T1c,
                   const double *T2c, double *R1p, double *R2p) {
  SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "residual: integrals not built");
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  const int64_t n2 = v * v * o * o;
  double *T1p = const_cast<double *>(T1c);
  double *T2p = const_cast<double *>(T2c);
  View T1 = V2(T1p, v, o), T2 = V4(T2p, v, v, o, o);
  View R1 = V2(R1p, v, o), R2 = V4(R2p, v, v, o, o);
  View V = V4(e.Vabij.get(), v, v, o, o), U = V4(e.Uabij.get(), v, v, o, o);
  View Vaibj = V4(e.Vaibj.get(), v, o, v, o), Vijkl = V4(e.Vijkl.get(), o, o, o, o),
       Vijka = V4(e.Vijka.get(), o, o, o, v);
  View Gij = V3(e.Gij.get(), G, o, o), Gai = V3(e.Gai.get(), G, v, o),
       Gab = V3(e.Gab.get(), G, v, v);
  zero(e, R1p, v * o);
  if (it == 0 && !initial_doubles) {  // MP2 start (cxx:47-51)
    copy(e, R2, V);
    return;
  }
  zero(e, R2p, n2);
  const int prec = e.precision;
  e.thaw();
  e.frozen = {T1p, T2p, e.Vabij.get(), e.Uabij.get(), e.Vaibj.get(), e.Vijka.get(),
              e.Gij.get(), e.Gai.get(), e.Gab.get()};

  DBuf Xb(&e.pool, n2);
  View X = V4(Xb.get(), v, v, o, o);
  add_t1t1(e, Xb.get(), T2p, T1p, 1.0, int(v), int(o));  // cxx:136-138
  e.frozen.push_back(Xb.get());

  DBuf Kacb(&e.pool, v * v), Lacb(&e.pool, v * v), Kkib(&e.pool, o * o), Lkib(&e.pool, o * o);
  View Kac = V2(Kacb.get(), v, v), Lac = V2(Lacb.get(), v, v), Kki = V2(Kkib.get(), o, o),
       Lki = V2(Lkib.get(), o, o);
  {
    // Kac = -2 V[cdkl] X[adkl] + V[dckl] X[adkl] = -U[cdkl] X[adkl] (cxx:146-147)
    ctr(e, -1.0, "cdkl,adkl->ac", U, X, 0.0, Kac, prec);
    copy(e, Lac, Kac);  // cxx:151
    // + 2 Gamma_Gca Gamma_Gdk T_dk (cxx:162-165)
    DBuf Zb(&e.pool, G);
    ctr(e, 1.0, "Gdk,dk->G", Gai, T1, 0.0, V1(Zb.get(), G), prec);
    ctr(e, 2.0, "Gca,G->ac", Gab, V1(Zb.get(), G), 1.0, Lac, prec);
    // - Gamma_Gck Gamma_Gda T_dk (cxx:166-169)
    DBuf Yb(&e.pool, G * v * o);
    ctr(e, 1.0, "Gda,dk->Gak", Gab, T1, 0.0, V3(Yb.get(), G, v, o), prec);
    ctr(e, -1.0, "Gck,Gak->ac", Gai, V3(Yb.get(), G, v, o), 1.0, Lac, prec);
    // Kki = 2 V[cdkl] X[cdil] - V[dckl] X[cdil] = U[cdkl] X[cdil] (cxx:172-173)
    ctr(e, 1.0, "cdkl,cdil->ki", U, X, 0.0, Kki, prec);
    copy(e, Lki, Kki);  // cxx:177
    ctr(e, 2.0, "klic,cl->ki", Vijka, T1, 1.0, Lki, prec);   // cxx:188
    ctr(e, -1.0, "lkic,cl->ki", Vijka, T1, 1.0, Lki, prec);  // cxx:189
    ctr(e, 1.0, "ac,cbij->abij", Lac, T2, 1.0, R2, prec);    // cxx:192
    ctr(e, -1.0, "ki,abkj->abij", Lki, T2, 1.0, R2, prec);   // cxx:195
  }
  {
    // dressed Gamma_ai and its terms (cxx:198-216)
    DBuf dGb(&e.pool, G * v * o), Qb(&e.pool, G * v * o);
    View dGai = V3(dGb.get(), G, v, o), Q = V3(Qb.get(), G, v, o);
    copy(e, dGai, Gai);
    ctr(e, -1.0, "Gki,ak->Gai", Gij, T1, 1.0, dGai, prec);   // cxx:205
    ctr(e, 1.0, "Gbc,cj->Gbj", Gab, T1, 0.0, Q, prec);       // cxx:208 (Gamma_Gbc T_cj)
    ctr(e, 1.0, "Gai,Gbj->abij", dGai, Q, 1.0, R2, prec);    // cxx:208-209
    ctr(e, -1.0, "jika,bk->abij", Vijka, T1, 1.0, R2, prec); // cxx:213
    DBuf Pb(&e.pool, v * o * o * o);
    View P = V4(Pb.get(), v, o, o, o);
    ctr(e, 1.0, "acik,cj->aikj", V, T1, 0.0, P, prec);       // cxx:214-215
    ctr(e, -1.0, "bk,aikj->abij", T1, P, 1.0, R2, prec);
  }
  {
    // Xakic (cxx:218-252)
    DBuf dGb(&e.pool, G * v * o);
    View dGai = V3(dGb.get(), G, v, o);
    copy(e, dGai, Gai);
    ctr(e, -1.0, "Gil,al->Gai", Gij, T1, 1.0, dGai, prec);  // cxx:227
    ctr(e, 1.0, "Gad,di->Gai", Gab, T1, 1.0, dGai, prec);   // cxx:230
    DBuf Xb2(&e.pool, v * o * o * v);
    View Xakic = V4(Xb2.get(), v, o, o, v);
    ctr(e, 1.0, "Gai,Gck->akic", dGai, Gai, 0.0, Xakic, prec);  // cxx:234
    DBuf Yb(&e.pool, n2);
    View Y = V4(Yb.get(), v, v, o, o);
    add_t1t1(e, Yb.get(), T2p, T1p, 2.0, int(v), int(o));    // cxx:238-240
    ctr(e, -0.5, "dclk,dail->akic", V, Y, 1.0, Xakic, prec);   // cxx:242
    ctr(e, 1.0, "dclk,adil->akic", V, T2, 1.0, Xakic, prec);   // cxx:243
    ctr(e, -0.5, "cdlk,adil->akic", V, T2, 1.0, Xakic, prec);  // cxx:245
    // Y = 2 T[cbkj] - T[bckj] (cxx:248-249)
    axpby(e, n2, 2.0, T2p, 0.0, Yb.get());
    sum(e, -1.0, T2, "bckj", 1.0, Y, "cbkj");
    ctr(e, 1.0, "akic,cbkj->abij", Xakic, Y, 1.0, R2, prec);  // cxx:251
  }
  {
    // Xakci (cxx:254-290)
    DBuf dGabb(&e.pool, G * v * v), dGijb(&e.pool, G * o * o);
    View dGab = V3(dGabb.get(), G, v, v), dGij = V3(dGijb.get(), G, o, o);
    copy(e, dGab, Gab);
    copy(e, dGij, Gij);
    ctr(e, -1.0, "Gcl,al->Gac", Gai, T1, 1.0, dGab, prec);  // cxx:270
    ctr(e, 1.0, "Gdk,di->Gki", Gai, T1, 1.0, dGij, prec);   // cxx:273
    DBuf Xb3(&e.pool, v * o * v * o);
    View Xakci = V4(Xb3.get(), v, o, v, o);
    ctr(e, 1.0, "Gac,Gki->akci", dGab, dGij, 0.0, Xakci, prec);  // cxx:277-278
    ctr(e, -0.5, "cdlk,dail->akci", V, T2, 1.0, Xakci, prec);    // cxx:284
    e.frozen.push_back(Xb3.get());
    ctr(e, -1.0, "akci,cbkj->abij", Xakci, T2, 1.0, R2, prec);   // cxx:288
    ctr(e, -1.0, "bkci,ackj->abij", Xakci, T2, 1.0, R2, prec);   // cxx:289
    e.unfreeze(Xb3.get());  // Xb3 returns to the pool at the end of this block
  }
  {
    // R_abij += R_baji (cxx:292)
    DBuf Rt(&e.pool, n2);
    copy(e, V4(Rt.get(), v, v, o, o), R2);
    sum(e, 1.0, V4(Rt.get(), v, v, o, o), "baji", 1.0, R2, "abij");
  }
  axpby(e, n2, 1.0, e.Vabij.get(), 1.0, R2p);  // cxx:301
  {
    // Xklij (cxx:303-334)
    DBuf Xb4(&e.pool, o * o * o * o);
    View Xklij = V4(Xb4.get(), o, o, o, o);
    copy(e, Xklij, Vijkl);
    ctr(e, 1.0, "klic,cj->klij", Vijka, T1, 1.0, Xklij, prec);  // cxx:308
    ctr(e, 1.0, "lkjc,ci->klij", Vijka, T1, 1.0, Xklij, prec);  // cxx:309
    ctr(e, 1.0, "klij,abkl->abij", Xklij, X, 1.0, R2, prec);    // cxx:313
    ctr(e, 1.0, "cdkl,cdij->klij", V, X, 0.0, Xklij, prec);     // cxx:317
    ctr(e, 1.0, "klij,abkl->abij", Xklij, T2, 1.0, R2, prec);   // cxx:330 (ppl = 1)
  }
  {
    // particle-particle ladder (cxx:374-453)
    if (slice_size <= 0) slice_size = int(o);
    DBuf dGabb(&e.pool, G * v * v);
    View dGab = V3(dGabb.get(), G, v, v);
    copy(e, dGab, Gab);
    ctr(e, -1.0, "Gbk,ak->Gab", Gai, T1, 1.0, dGab, prec);  // cxx:397-400
    const int nslices = int((v + slice_size - 1) / slice_size);
    std::vector<DBuf> sl(nslices);
    std::vector<int64_t> lo(nslices), hi(nslices);
    for (int s = 0; s < nslices; ++s) {  // cxx:402-416
      lo[s] = int64_t(s) * slice_size;
      hi[s] = std::min<int64_t>(lo[s] + slice_size, v);
      sl[s] = DBuf(&e.pool, G * (hi[s] - lo[s]) * v);
      copy(e, V3(sl[s].get(), G, hi[s] - lo[s], v), subview(dGab, {0, lo[s], 0}, {G, hi[s], v}));
    }
    const int64_t smax = std::min<int64_t>(slice_size, v);
    DBuf Vxy(&e.pool, smax * smax * v * v), Rxy(&e.pool, smax * smax * o * o);
    for (int m = 0; m < nslices; ++m)
      for (int n = m; n < nslices; ++n) {  // cxx:418-451
        const int64_t Nx = hi[n] - lo[n], Ny = hi[m] - lo[m];
        View Vxycd = V4(Vxy.get(), Nx, Ny, v, v);
        View Rxyij = V4(Rxy.get(), Nx, Ny, o, o);
        ctr(e, 1.0, "Gxc,Gyd->xycd", V3(sl[n].get(), G, Nx, v), V3(sl[m].get(), G, Ny, v), 0.0,
            Vxycd, e.prec_ladder);  // cxx:431-434
        ctr(e, 1.0, "xycd,cdij->xyij", Vxycd, X, 0.0, Rxyij, e.prec_ladder);  // cxx:443
        // sliceIntoResiduum (ClusterSinglesDoublesAlgorithm.cxx:887-915)
        const int64_t a = lo[n], b = lo[m];
        copy(e, subview(R2, {a, b, 0, 0}, {a + Nx, b + Ny, o, o}), Rxyij, 1.0, 1.0);
        if (a > b) sum(e, 1.0, Rxyij, "xyij", 1.0, subview(R2, {b, a, 0, 0}, {b + Ny, a + Nx, o, o}), "yxji");
      }
  }
  {
    // T1 equations (cxx:457-492)
    ctr(e, 1.0, "ac,ci->ai", Kac, T1, 1.0, R1, prec);    // cxx:466
    ctr(e, -1.0, "ki,ak->ai", Kki, T1, 1.0, R1, prec);   // cxx:467
    DBuf Kckb(&e.pool, v * o), Skib(&e.pool, o * o);
    View Kck = V2(Kckb.get(), v, o);
    ctr(e, 1.0, "cdkl,dl->ck", U, T1, 0.0, Kck, prec);   // cxx:470-471
    ctr(e, 2.0, "ck,caki->ai", Kck, T2, 1.0, R1, prec);  // cxx:474
    ctr(e, -1.0, "ck,caik->ai", Kck, T2, 1.0, R1, prec); // cxx:475
    ctr(e, 1.0, "ck,ci->ki", Kck, T1, 0.0, V2(Skib.get(), o, o), prec);  // cxx:476
    ctr(e, 1.0, "ak,ki->ai", T1, V2(Skib.get(), o, o), 1.0, R1, prec);
    ctr(e, 2.0, "acik,ck->ai", V, T1, 1.0, R1, prec);      // cxx:477
    ctr(e, -1.0, "ciak,ck->ai", Vaibj, T1, 1.0, R1, prec); // cxx:478
    DBuf Ygb(&e.pool, G * v * o);
    View Yg = V3(Ygb.get(), G, v, o);
    ctr(e, 1.0, "Gdk,cdik->Gci", Gai, X, 0.0, Yg, prec);  // cxx:480-481
    ctr(e, 2.0, "Gca,Gci->ai", Gab, Yg, 1.0, R1, prec);
    ctr(e, 1.0, "Gck,cdik->Gdi", Gai, X, 0.0, Yg, prec);  // cxx:484-485
    ctr(e, -1.0, "Gda,Gdi->ai", Gab, Yg, 1.0, R1, prec);
    ctr(e, -2.0, "klic,ackl->ai", Vijka, X, 1.0, R1, prec);  // cxx:489
    ctr(e, 1.0, "lkic,ackl->ai", Vijka, X, 1.0, R1, prec);   // cxx:490
  }
  e.thaw();
}

void ccsd_energy(Engine &e, const double *T1, const double *T2, double &E, double &direct,
                 double &exchange) {
  SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "energy: integrals not built");
  energy_terms(e, T1, T2, direct, exchange);
  E = direct + exchange;
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------

namespace {

// Solve the bordered DIIS system (dsysv_ in DiisMixer.cxx:16-40) in FP64 by
// Gaussian elimination with partial pivoting.
std::vector<double> solve_small(std::vector<std::vector<double>> A, std::vector<double> b) {
  const int n = int(b.size());
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    SISI_REQUIRE(std::isfinite(A[p][c]), SISI_ERR_NUMERIC,
                 "DIIS: non-finite overlap (the iteration diverged)");
    SISI_REQUIRE(A[p][c] != 0.0, SISI_ERR_NUMERIC, "DIIS: singular overlap matrix");
    std::swap(A[p], A[c]);
    std::swap(b[p], b[c]);
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int k = c; k < n; ++k) A[r][k] -= f * A[c][k];
      b[r] -= f * b[c];
    }
  }
  std::vector<double> x(n);
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= A[r][k] * x[k];
    x[r] = s / A[r][r];
  }
  return x;
}

double fock_dot(Engine &e, const double *x1, const double *x2, const double *y1,
                const double *y2) {
  return dot(e, int64_t(e.n1()), x1, y1) + dot(e, int64_t(e.n2()), x2, y2);
}

}  // namespace

void ccsd_begin(Engine &e, const sisi_ccsd_opts &opts, const double *T1h, const double *T2h) {
  SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "ccsd: integrals not built");
  SISI_REQUIRE(opts.mixer == SISI_MIXER_LINEAR || opts.mixer == SISI_MIXER_DIIS,
               SISI_ERR_INVALID, "Mixer not implemented");
  SISI_REQUIRE(opts.mixer != SISI_MIXER_DIIS || opts.max_residua >= 1, SISI_ERR_INVALID,
               "maxResidua must be >= 1");
  e.opts = opts;
  e.T1 = DBuf(&e.pool, e.n1());
  e.T2 = DBuf(&e.pool, e.n2());
  e.run_initial = (T2h != nullptr);
  if (T1h) e.upload(e.T1, T1h, e.n1());
  else zero(e, e.T1.get(), int64_t(e.n1()));
  if (T2h) e.upload(e.T2, T2h, e.n2());
  else zero(e, e.T2.get(), int64_t(e.n2()));
  e.iteration = 0;
  e.fp64_iterations = 0;
  e.prev_energy = 0.0;
  e.last_energy = e.last_direct = e.last_exchange = 0.0;
  e.last_delta = 1.0;
  e.converged = false;
  e.energies.clear();
  e.diis.clear();
  e.B.clear();
  e.diis_next = e.diis_count = 0;
  e.lin_have = false;
  e.lin1.reset();
  e.lin2.reset();
  if (opts.mixer == SISI_MIXER_DIIS) {
    const int N = opts.max_residua;
    e.diis.resize(N);
    e.B.assign(N + 1, std::vector<double>(N + 1, 0.0));
    for (int k = 1; k <= N; ++k) e.B[0][k] = e.B[k][0] = -1.0;  // DiisMixer.cxx:83-96
  }
  e.cur_precision = opts.precision;
  e.run_active = true;
}

void ccsd_iterate(Engine &e, int n) {
  SISI_REQUIRE(e.run_active, SISI_ERR_STATE, "ccsd: no run (call begin first)");
  const sisi_ccsd_opts &o = e.opts;
  for (int step = 0; step < n && !e.converged && e.iteration < o.max_iterations; ++step) {
    // precision schedule: bulk precision until |dT|/|T| <= switch_threshold,
    // FP64 afterwards (and FP64 throughout for the plain path)
    const int prec = (o.precision != SISI_FP64 && e.cur_precision != SISI_FP64) ? o.precision
                                                                                : SISI_FP64;
    set_precision(e, prec);
    DBuf R1(&e.pool, e.n1()), R2(&e.pool, e.n2());
    ccsd_residual(e, e.iteration, e.run_initial, o.slice_size, e.T1.get(), e.T2.get(), R1.get(),
                  R2.get());
    DBuf d1(&e.pool, e.n1()), d2(&e.pool, e.n2());
    estimate_amplitudes(e, 2, R1.get(), e.T1.get(), d1.get(), o.level_shift);
    estimate_amplitudes(e, 4, R2.get(), e.T2.get(), d2.get(), o.level_shift);
    if (o.mixer == SISI_MIXER_DIIS) {
      // DiisMixer::append (DiisMixer.cxx:103-180)
      const int N = o.max_residua;
      DiisSlot &s = e.diis[e.diis_next];
      s.t1 = std::move(R1);
      s.t2 = std::move(R2);
      s.d1 = std::move(d1);
      s.d2 = std::move(d2);
      for (int i = 0; i < N; ++i) {
        if (!e.diis[i].d2.get()) continue;
        const double ov = 2.0 * fock_dot(e, e.diis[i].d1.get(), e.diis[i].d2.get(), s.d1.get(),
                                         s.d2.get());
        e.B[e.diis_next + 1][i + 1] = ov;
        e.B[i + 1][e.diis_next + 1] = ov;
      }
      if (e.diis_count < N) ++e.diis_count;
      const int dim = e.diis_count + 1;
      std::vector<std::vector<double>> Bm(dim, std::vector<double>(dim));
      for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) Bm[r][c] = e.B[r][c];
      std::vector<double> rhs(dim, 0.0);
      rhs[0] = -1.0;
      std::vector<double> col = solve_small(Bm, rhs);
      DBuf n1(&e.pool, e.n1()), n2(&e.pool, e.n2());
      zero(e, n1.get(), int64_t(e.n1()));
      zero(e, n2.get(), int64_t(e.n2()));
      for (int j = 0; j < e.diis_count; ++j) {
        const int i = (e.diis_next + N - j) % N;
        axpby(e, int64_t(e.n1()), col[i + 1], e.diis[i].t1.get(), 1.0, n1.get());
        axpby(e, int64_t(e.n2()), col[i + 1], e.diis[i].t2.get(), 1.0, n2.get());
      }
      e.diis_next = (e.diis_next + 1) % N;
      e.T1 = std::move(n1);
      e.T2 = std::move(n2);
      // amplitude change of this iteration, read back for the test below
      const DiisSlot &cur = e.diis[(e.diis_next + N - 1) % N];
      const double dd = fock_dot(e, cur.d1.get(), cur.d2.get(), cur.d1.get(), cur.d2.get());
      const double tt = fock_dot(e, e.T1.get(), e.T2.get(), e.T1.get(), e.T2.get());
      e.last_delta = std::sqrt(std::fabs(dd / tt));
    } else {
      // LinearMixer::append (LinearMixer.cxx:30-44)
      if (e.lin_have) {
        axpby(e, int64_t(e.n1()), 1.0 - o.mixing_ratio, e.lin1.get(), o.mixing_ratio, R1.get());
        axpby(e, int64_t(e.n2()), 1.0 - o.mixing_ratio, e.lin2.get(), o.mixing_ratio, R2.get());
      }
      e.lin1 = std::move(R1);
      e.lin2 = std::move(R2);
      e.lin_have = true;
      DBuf n1(&e.pool, e.n1()), n2(&e.pool, e.n2());
      copy(e, V1(n1.get(), int64_t(e.n1())), V1(e.lin1.get(), int64_t(e.n1())));
      copy(e, V1(n2.get(), int64_t(e.n2())), V1(e.lin2.get(), int64_t(e.n2())));
      e.T1 = std::move(n1);
      e.T2 = std::move(n2);
      const double dd = fock_dot(e, d1.get(), d2.get(), d1.get(), d2.get());
      const double tt = fock_dot(e, e.T1.get(), e.T2.get(), e.T1.get(), e.T2.get());
      e.last_delta = std::sqrt(std::fabs(dd / tt));
    }
    double E, dir, exc;
    ccsd_energy(e, e.T1.get(), e.T2.get(), E, dir, exc);
    SISI_REQUIRE(std::isfinite(E) && std::isfinite(e.last_delta), SISI_ERR_NUMERIC,
                 "CCSD iteration " + std::to_string(e.iteration + 1) +
                     " diverged (non-finite energy or amplitudes)");
    e.last_energy = E;
    e.last_direct = dir;
    e.last_exchange = exc;
    e.energies.push_back(E);
    if (prec == SISI_FP64) e.fp64_iterations++;
    e.iteration++;
    // hand over to FP64 once the bulk iterations reached the threshold
    if (prec != SISI_FP64 && e.last_delta <= o.switch_threshold && e.iteration > 1)
      e.cur_precision = SISI_FP64;
    // convergence test (ClusterSinglesDoublesAlgorithm.cxx:103-106), FP64 only
    const double de = std::fabs((E - e.prev_energy) / E);
    if (prec == SISI_FP64 && de < std::fabs(o.energy_convergence) &&
        e.last_delta * e.last_delta < std::fabs(o.amplitudes_convergence * o.amplitudes_convergence))
      e.converged = true;
    e.prev_energy = E;
  }
}

void fill_report(const Engine &e, sisi_ccsd_report *r) {
  r->energy = e.last_energy;
  r->direct = e.last_direct;
  r->exchange = e.last_exchange;
  r->delta = e.last_delta;
  r->iterations = e.iteration;
  r->converged = e.converged ? 1 : 0;
  r->fp64_iterations = e.fp64_iterations;
  r->precision = e.cur_precision;
  r->status = SISI_OK;
}

}  // namespace sisi
