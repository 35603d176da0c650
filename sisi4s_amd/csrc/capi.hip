// extern "C" boundary: argument checks, exception -> status conversion,
// host <-> device staging.  See include/sisi_engine.h for the contract.
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ccsd.hpp"
#include "contract.hpp"
#include "engine.hpp"
#include "gemm.hpp"
#include "tune.hpp"
#include "tensor_ops.hpp"
#include "triples.hpp"

extern "C" int sisi_generate_ueg_impl(int No, int Nv, double rs, int hartree_fock, double madelung,
                                      int *NG, double *gamma, double *eps_i, double *eps_a,
                                      const char **error);
extern "C" int sisi_generate_system_impl(uint64_t seed, int No, int Nv, int NG, double *gamma,
                                         double *eps_i, double *eps_a);

namespace sisi {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

template <typename F>
static int guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return SISI_OK;
  } catch (const Error &err) {
    g_last_error = err.what();
    return err.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return SISI_ERR_ALLOC;
  } catch (const std::exception &ex) {
    g_last_error = ex.what();
    return SISI_ERR_INVALID;
  } catch (...) {
    g_last_error = "unknown error";
    return SISI_ERR_INVALID;
  }
}

static void check_finite(const double *p, size_t n, const char *what) {
  for (size_t i = 0; i < n; ++i)
    SISI_REQUIRE(std::isfinite(p[i]), SISI_ERR_NUMERIC, std::string(what) + " has NaN/Inf");
}

}  // namespace sisi

using namespace sisi;

struct sisi_engine {
  Engine e;
  explicit sisi_engine(const sisi_dims &d) : e(d) {}
};

#define ENGINE_REQUIRED(h) SISI_REQUIRE((h) != nullptr, SISI_ERR_INVALID, "null engine handle")

extern "C" {

const char *sisi_last_error(void) { return g_last_error.c_str(); }

void sisi_ccsd_opts_default(sisi_ccsd_opts *o) {
  if (!o) return;
  o->mixer = SISI_MIXER_LINEAR;
  o->max_residua = 4;
  o->mixing_ratio = 1.0;
  o->max_iterations = 16;
  o->amplitudes_convergence = 1e-5;
  o->energy_convergence = 1e-6;
  o->level_shift = 0.0;
  o->slice_size = -1;
  o->precision = SISI_FP64;
  o->switch_threshold = 1e-5;
  o->refine_interval = 0;
  o->refine_precision = SISI_BF16;
  o->ppl = 1;
  o->distinguishable = 0;
  o->coarse_threshold = 0.0;
  o->coarse_precision = SISI_BF16;
  o->graphs = 0;
  o->bulk_refine_interval = 0;
}

int sisi_engine_create(const sisi_dims *dims, sisi_engine **out) {
  return guarded([&] {
    SISI_REQUIRE(dims != nullptr && out != nullptr, SISI_ERR_INVALID, "null argument");
    *out = nullptr;
    *out = new sisi_engine(*dims);
  });
}

int sisi_engine_destroy(sisi_engine *h) {
  return guarded([&] { delete h; });
}

int sisi_engine_set_vertex(sisi_engine *h, const double *gamma) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    const size_t n = size_t(e.NG) * e.Np * e.Np;
    SISI_REQUIRE(gamma != nullptr, SISI_ERR_INVALID, "null CoulombVertex");
    check_finite(gamma, n, "CoulombVertex");
    e.cplx = false;
    e.upload(e.gamma, gamma, n);
    e.have_vertex = true;
    e.have_integrals = false;
    e.run_active = false;
    e.drop_triples_cache();
    e.drop_residual_graph();
  });
}

int sisi_engine_set_vertex_complex(sisi_engine *h, const double *gamma) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    const size_t n = 2 * size_t(e.NG) * e.Np * e.Np;
    SISI_REQUIRE(gamma != nullptr, SISI_ERR_INVALID, "null CoulombVertex");
    check_finite(gamma, n, "CoulombVertex");
    e.cplx = true;
    e.upload_t(e.gamma, gamma, n);
    e.have_vertex = true;
    e.have_integrals = false;
    e.run_active = false;
    e.drop_triples_cache();
    e.drop_residual_graph();
  });
}

int sisi_engine_set_eigenenergies(sisi_engine *h, const double *eps_i, const double *eps_a) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(eps_i && eps_a, SISI_ERR_INVALID, "null eigenenergies");
    check_finite(eps_i, size_t(e.No), "HoleEigenEnergies");
    check_finite(eps_a, size_t(e.Nv), "ParticleEigenEnergies");
    e.upload(e.eps_i, eps_i, size_t(e.No));
    e.upload(e.eps_a, eps_a, size_t(e.Nv));
    e.have_eps = true;
    e.drop_eom();  // F_mi and F_ae hold the eigenenergies
  });
}

int sisi_integrals_from_vertex(sisi_engine *h) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    integrals_from_vertex(h->e);
  });
}

int sisi_get_integrals(sisi_engine *h, int which, double *out) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "integrals not built");
    const size_t o = size_t(e.No), v = size_t(e.Nv);
    if (e.cplx) {  // complex128, interleaved
      switch (which) {
        case SISI_PPHH: e.download_t(out, e.Vabij, 2 * v * v * o * o); break;
        case SISI_PHPH: e.download_t(out, e.Vaibj, 2 * v * o * v * o); break;
        case SISI_HHHH: e.download_t(out, e.Vijkl, 2 * o * o * o * o); break;
        case SISI_HHHP: e.download_t(out, e.Vijka, 2 * o * o * o * v); break;
        case SISI_HHPP: e.download_t(out, e.Vijab, 2 * o * o * v * v); break;
        case SISI_PHHP: e.download_t(out, e.Vaijb, 2 * v * o * o * v); break;
        case SISI_PHHH: e.download_t(out, e.Vaijk, 2 * v * o * o * o); break;
        default: throw Error(SISI_ERR_INVALID, "unknown integral class");
      }
      return;
    }
    switch (which) {
      case SISI_PPHH: e.download(out, e.Vabij, v * v * o * o); break;
      case SISI_PHPH: e.download(out, e.Vaibj, v * o * v * o); break;
      case SISI_HHHH: e.download(out, e.Vijkl, o * o * o * o); break;
      case SISI_HHHP: e.download(out, e.Vijka, o * o * o * v); break;
      case SISI_PPPP:
        if (!e.Vabcd.get()) build_vabcd(e);
        e.download(out, e.Vabcd, v * v * v * v);
        break;
      case SISI_PPPH:
        if (!e.Vabci.get()) build_vabcd(e);
        e.download(out, e.Vabci, v * v * v * o);
        break;
      case SISI_HHPP: {
        DBuf t(&e.pool, o * o * v * v);
        sum(e, 1.0, dense(e.Vabij.get(), {int64_t(v), int64_t(v), int64_t(o), int64_t(o)}), "abij",
            0.0, dense(t.get(), {int64_t(o), int64_t(o), int64_t(v), int64_t(v)}), "ijab");
        e.download(out, t, o * o * v * v);
        break;
      }
      case SISI_PHHP: {  // V_aijb = V_abji (CoulombIntegralsFromVertex.cxx:515)
        DBuf t(&e.pool, v * o * o * v);
        sum(e, 1.0, dense(e.Vabij.get(), {int64_t(v), int64_t(v), int64_t(o), int64_t(o)}), "abji",
            0.0, dense(t.get(), {int64_t(v), int64_t(o), int64_t(o), int64_t(v)}), "aijb");
        e.download(out, t, v * o * o * v);
        break;
      }
      default: throw Error(SISI_ERR_INVALID, "unknown integral class");
    }
  });
}

int sisi_engine_set_precision(sisi_engine *h, int precision) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    set_precision(h->e, precision);
  });
}

int sisi_engine_set_coulomb_factors(sisi_engine *h, int NR, int factors_slice_size,
                                    const double *lambda, const double *pi,
                                    const double *pi_out) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    SISI_REQUIRE(NR >= 0, SISI_ERR_INVALID, "negative NR");
    set_coulomb_factors(h->e, NR, factors_slice_size, lambda, pi, pi_out);
  });
}

int sisi_set_bf16_backend(int library) {
  return guarded([&] {
    SISI_REQUIRE(library == 0 || library == 1, SISI_ERR_INVALID, "backend must be 0 or 1");
    set_b3_backend(library);
  });
}

int sisi_gemm_set_variant(int passes, int variant) {
  return guarded([&] {
    SISI_REQUIRE(passes == 1 || passes == 3, SISI_ERR_INVALID, "passes must be 1 or 3");
    set_b3_variant(passes, variant);
  });
}

int sisi_gemm_variant_count(int passes) {
  return passes == 1 || passes == 3 ? b3_variant_count(passes) : 0;
}

const char *sisi_gemm_variant_name(int passes, int variant) {
  return passes == 1 || passes == 3 ? b3_variant_name(passes, variant) : "";
}

int sisi_tune_enable(int on) {
  return guarded([&] { tune_set_enabled(on != 0); });
}

int sisi_tune_dump(char *buf, int len, int *needed) {
  return guarded([&] {
    std::string all = "# cache: " + tune_cache_path() + "\n";
    for (const auto &l : tune_dump()) all += l + "\n";
    if (needed) *needed = int(all.size()) + 1;
    if (buf && len > 0) {
      const size_t n = std::min<size_t>(all.size(), size_t(len - 1));
      std::memcpy(buf, all.data(), n);
      buf[n] = 0;
    }
  });
}

int sisi_engine_set_integrals(sisi_engine *h, int which, const double *V) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    set_integrals(h->e, which, V);
  });
}

int sisi_engine_set_formulation(sisi_engine *h, int formulation) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    SISI_REQUIRE(formulation == SISI_FORMULATION_VERTEX || formulation == SISI_FORMULATION_INTEGRALS ||
                     formulation == SISI_FORMULATION_ONLY_PPL ||
                     formulation == SISI_FORMULATION_DRCCD || formulation == SISI_FORMULATION_UCCSD,
                 SISI_ERR_INVALID, "unknown formulation");
    SISI_REQUIRE(formulation == SISI_FORMULATION_VERTEX || formulation == SISI_FORMULATION_DRCCD ||
                     !h->e.cplx,
                 SISI_ERR_INVALID, "this formulation is built for the real path");
    h->e.formulation = formulation;
  });
}

int sisi_eom_set_amplitudes(sisi_engine *h, const double *T1, const double *T2) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    SISI_REQUIRE(T1 && T2, SISI_ERR_INVALID, "null amplitudes");
    SISI_REQUIRE(!h->e.cplx, SISI_ERR_INVALID, "the EOM right-applies are built for real amplitudes");
    eom_set_amplitudes(h->e, T1, T2);
  });
}

int sisi_eom_ip_sigma(sisi_engine *h, const double *T1, const double *T2, const double *r1,
                      const double *r2, double *s1, double *s2) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(r1 && r2 && s1 && s2, SISI_ERR_INVALID, "null argument");
    SISI_REQUIRE((T1 == nullptr) == (T2 == nullptr), SISI_ERR_INVALID,
                 "T1 and T2 are given together (or both null: the amplitudes set)");
    SISI_REQUIRE(!e.cplx, SISI_ERR_INVALID, "the IP right-apply is built for real amplitudes");
    const size_t o = size_t(e.No), v = size_t(e.Nv);
    if (T1) eom_set_amplitudes(e, T1, T2);
    DBuf q1, q2, w1(&e.pool, o), w2(&e.pool, o * o * v);
    e.upload(q1, r1, o);
    e.upload(q2, r2, o * o * v);
    eom_ip_sigma(e, q1.get(), q2.get(), w1.get(), w2.get());
    e.download(s1, w1, o);
    e.download(s2, w2, o * o * v);
  });
}

int sisi_eom_ea_sigma(sisi_engine *h, const double *T1, const double *T2, const double *r1,
                      const double *r2, double *s1, double *s2) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(r1 && r2 && s1 && s2, SISI_ERR_INVALID, "null argument");
    SISI_REQUIRE((T1 == nullptr) == (T2 == nullptr), SISI_ERR_INVALID,
                 "T1 and T2 are given together (or both null: the amplitudes set)");
    SISI_REQUIRE(!e.cplx, SISI_ERR_INVALID, "the EA right-apply is built for real amplitudes");
    const size_t o = size_t(e.No), v = size_t(e.Nv);
    if (T1) eom_set_amplitudes(e, T1, T2);
    DBuf q1, q2, w1(&e.pool, v), w2(&e.pool, o * v * v);
    e.upload(q1, r1, v);
    e.upload(q2, r2, o * v * v);
    eom_ea_sigma(e, q1.get(), q2.get(), w1.get(), w2.get());
    e.download(s1, w1, v);
    e.download(s2, w2, o * v * v);
  });
}

int sisi_engine_set_drccd(sisi_engine *h, int unrestricted, int linearized,
                          int adjacent_pairs_exchange) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    set_drccd(h->e, unrestricted, linearized, adjacent_pairs_exchange);
  });
}

int sisi_engine_set_energy_variant(sisi_engine *h, int unrestricted, int antisymmetrize) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    h->e.spins = unrestricted ? 1.0 : 2.0;  // ClusterSinglesDoublesAlgorithm.cxx:133
    h->e.anti_energy = antisymmetrize != 0;  // cxx:134, 163-167
  });
}

int sisi_engine_set_hp_fock(sisi_engine *h, const double *fia) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    if (fia == nullptr) {
      e.fai = DBuf();
      e.noncanonical = 0.0;
      return;
    }
    SISI_REQUIRE(!e.cplx, SISI_ERR_INVALID, "HPFockMatrix: built for a real vertex");
    const int64_t o = e.No, v = e.Nv, w = 1;
    check_finite(fia, size_t(w * o * v), "HPFockMatrix");
    // f[i,a] at i + No a (complex128 interleaved in complex mode) -> f[a,i]
    std::vector<double> t(size_t(w * o * v));
    for (int64_t i = 0; i < o; ++i)
      for (int64_t a = 0; a < v; ++a)
        for (int64_t c = 0; c < w; ++c) t[size_t(w * (a + v * i) + c)] = fia[w * (i + o * a) + c];
    e.upload_t(e.fai, t.data(), t.size());
    e.fai_cplx = e.cplx;
  });
}

int sisi_engine_get_noncanonical_energy(sisi_engine *h, double *out) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    SISI_REQUIRE(out != nullptr, SISI_ERR_INVALID, "null output");
    *out = h->e.noncanonical;
  });
}

int sisi_engine_set_residual_variant(sisi_engine *h, int ppl, int distinguishable) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    h->e.ppl = ppl != 0;
    h->e.distinguishable = distinguishable != 0;
  });
}

int sisi_ccsd_residual(sisi_engine *h, int iteration, int initial_doubles, int slice_size,
                       const double *T1, const double *T2, double *R1, double *R2) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(T1 && T2 && R1 && R2, SISI_ERR_INVALID, "null amplitude/residual pointer");
    SISI_REQUIRE(iteration >= 0, SISI_ERR_INVALID, "negative iteration");
    DBuf t1, t2, r1(&e.pool, e.n1()), r2(&e.pool, e.n2());
    e.upload_t(t1, T1, e.n1());
    e.upload_t(t2, T2, e.n2());
    ccsd_residual(e, iteration, initial_doubles != 0, slice_size, t1.get(), t2.get(), r1.get(),
                  r2.get());
    e.download_t(R1, r1, e.n1());
    e.download_t(R2, r2, e.n2());
  });
}

static int residual_variant(sisi_engine *h, bool anchor, int slice_size, const double *T1,
                            const double *T2, double *R1, double *R2) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(T1 && T2 && R1 && R2, SISI_ERR_INVALID, "null amplitude/residual pointer");
    SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "integrals not built");
    DBuf t1, t2, r1(&e.pool, e.n1()), r2(&e.pool, e.n2());
    e.upload(t1, T1, e.n1());
    e.upload(t2, T2, e.n2());
    if (anchor) ccsd_residual_anchor(e, slice_size, t1.get(), t2.get(), r1.get(), r2.get());
    else ccsd_residual_delta(e, slice_size, t1.get(), t2.get(), r1.get(), r2.get());
    e.download(R1, r1, e.n1());
    e.download(R2, r2, e.n2());
  });
}

int sisi_ccsd_residual_anchor(sisi_engine *h, int slice_size, const double *T1,
                              const double *T2, double *R1, double *R2) {
  return residual_variant(h, true, slice_size, T1, T2, R1, R2);
}

int sisi_ccsd_residual_refine(sisi_engine *h, int slice_size, const double *T1,
                              const double *T2, double *R1, double *R2) {
  return residual_variant(h, false, slice_size, T1, T2, R1, R2);
}

int sisi_ccsd_energy(sisi_engine *h, const double *T1, const double *T2, double out[3]) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(T1 && T2 && out, SISI_ERR_INVALID, "null pointer");
    DBuf t1, t2;
    e.upload_t(t1, T1, e.n1());
    e.upload_t(t2, T2, e.n2());
    ccsd_energy(e, t1.get(), t2.get(), out[0], out[1], out[2]);
  });
}

int sisi_mp2_energy(sisi_engine *h, double *energy) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(energy, SISI_ERR_INVALID, "null pointer");
    DBuf t1(&e.pool, e.n1()), t2(&e.pool, e.n2()), r1(&e.pool, e.n1()), r2(&e.pool, e.n2()),
        d1(&e.pool, e.n1()), d2(&e.pool, e.n2());
    zero(e, t1.get(), int64_t(e.n1()));
    zero(e, t2.get(), int64_t(e.n2()));
    ccsd_residual(e, 0, false, -1, t1.get(), t2.get(), r1.get(), r2.get());
    estimate_amplitudes(e, 2, r1.get(), t1.get(), d1.get(), 0.0);
    estimate_amplitudes(e, 4, r2.get(), t2.get(), d2.get(), 0.0);
    double dir, exc;
    ccsd_energy(e, r1.get(), r2.get(), *energy, dir, exc);
  });
}

int sisi_ccsd_begin(sisi_engine *h, const sisi_ccsd_opts *opts, const double *T1, const double *T2) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    sisi_ccsd_opts o;
    if (opts) o = *opts;
    else sisi_ccsd_opts_default(&o);
    ccsd_begin(h->e, o, T1, T2);
  });
}

int sisi_ccsd_iterate(sisi_engine *h, int n, sisi_ccsd_report *report) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    SISI_REQUIRE(n >= 0, SISI_ERR_INVALID, "negative iteration count");
    ccsd_iterate(h->e, n);
    if (report) fill_report(h->e, report);
  });
}

int sisi_ccsd_solve(sisi_engine *h, const sisi_ccsd_opts *opts, const double *T1,
                    const double *T2, sisi_ccsd_report *report) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    sisi_ccsd_opts o;
    if (opts) o = *opts;
    else sisi_ccsd_opts_default(&o);
    ccsd_begin(h->e, o, T1, T2);
    ccsd_iterate(h->e, o.max_iterations);
    Engine &e = h->e;
    if (o.max_iterations == 0) {  // energy of the given amplitudes (cxx:116-119)
      double E, d, x;
      ccsd_energy(e, e.T1.get(), e.T2.get(), E, d, x, &e.last_energy_imag);
      e.last_energy = E;
      e.last_direct = d;
      e.last_exchange = x;
    }
    if (report) fill_report(e, report);
  });
}

int sisi_ccsd_get_amplitudes(sisi_engine *h, double *T1, double *T2) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(e.run_active, SISI_ERR_STATE, "no CCSD run");
    if (T1) e.download_t(T1, e.T1, e.n1());
    if (T2) e.download_t(T2, e.T2, e.n2());
  });
}

int sisi_ccsd_pair_energies(sisi_engine *h, double *pair) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(pair, SISI_ERR_INVALID, "null pointer");
    SISI_REQUIRE(e.run_active, SISI_ERR_STATE, "no CCSD run");
    const size_t n = size_t(e.No) * e.No * (e.cplx ? 2 : 1);
    DBuf p(&e.pool, n);
    ccsd_pair_energies(e, e.T1.get(), e.T2.get(), p.get());
    e.download_t(pair, p, n);
  });
}

int sisi_ccsd_get_energies(sisi_engine *h, double *energies, int max, int *n) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    const int cnt = int(e.energies.size());
    if (n) *n = cnt;
    if (energies)
      for (int k = 0; k < cnt && k < max; ++k) energies[k] = e.energies[k];
  });
}

static int triples_common(sisi_engine *h, const double *T1, const double *T2, int precision,
                          const int *blocks, int nblocks, double *energy) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    Engine &e = h->e;
    SISI_REQUIRE(energy, SISI_ERR_INVALID, "null pointer");
    SISI_REQUIRE(precision == SISI_FP64 || precision == SISI_FP32 || precision == SISI_BF16X3,
                 SISI_ERR_INVALID, "unknown precision");
    SISI_REQUIRE((T1 == nullptr) == (T2 == nullptr), SISI_ERR_INVALID,
                 "give both amplitudes or neither");
    DBuf t1, t2;
    const double *d1, *d2;
    if (T1) {
      e.upload_t(t1, T1, e.n1());
      e.upload_t(t2, T2, e.n2());
      d1 = t1.get();
      d2 = t2.get();
    } else {
      SISI_REQUIRE(e.run_active, SISI_ERR_STATE, "no amplitudes: run CCSD or pass T1/T2");
      d1 = e.T1.get();
      d2 = e.T2.get();
    }
    std::vector<std::array<int, 3>> bl;
    if (blocks) {
      SISI_REQUIRE(nblocks > 0, SISI_ERR_INVALID, "empty block list");
      for (int b = 0; b < nblocks; ++b) bl.push_back({blocks[3 * b], blocks[3 * b + 1], blocks[3 * b + 2]});
    }
    *energy = triples_energy(e, d1, d2, precision, bl);
  });
}

int sisi_triples_energy(sisi_engine *h, const double *T1, const double *T2, int precision,
                        double *energy) {
  return triples_common(h, T1, T2, precision, nullptr, 0, energy);
}

int sisi_triples_energy_blocks(sisi_engine *h, const double *T1, const double *T2, int precision,
                               const int *blocks, int nblocks, double *energy) {
  if (!blocks) {
    g_last_error = "null block list";
    return SISI_ERR_INVALID;
  }
  return triples_common(h, T1, T2, precision, blocks, nblocks, energy);
}

int sisi_ccsd_solve_batched(int nsys, const sisi_dims *dims, const double *const *gammas,
                            const double *const *eps_i, const double *const *eps_a,
                            const sisi_ccsd_opts *opts, int workers, double *energies,
                            sisi_ccsd_report *reports) {
  return guarded([&] {
    SISI_REQUIRE(nsys >= 0 && dims && gammas && eps_i && eps_a && energies, SISI_ERR_INVALID,
                 "null argument");
    sisi_ccsd_opts o;
    if (opts) o = *opts;
    else sisi_ccsd_opts_default(&o);
    if (workers < 1) workers = 1;
    workers = std::min(workers, std::max(nsys, 1));
    int device = 0;
    SISI_HIP(hipGetDevice(&device));
    std::atomic<int> next(0);
    std::vector<int> status(size_t(nsys), SISI_OK);
    std::vector<std::string> msg(static_cast<size_t>(nsys));
    auto worker = [&]() {
      (void)hipSetDevice(device);
      for (int s = next++; s < nsys; s = next++) {
        const int rc = guarded([&] {
          Engine e(dims[s]);
          const size_t np = size_t(dims[s].No) + dims[s].Nv;
          SISI_REQUIRE(gammas[s] && eps_i[s] && eps_a[s], SISI_ERR_INVALID, "null system input");
          check_finite(gammas[s], size_t(dims[s].NG) * np * np, "CoulombVertex");
          check_finite(eps_i[s], size_t(dims[s].No), "HoleEigenEnergies");
          check_finite(eps_a[s], size_t(dims[s].Nv), "ParticleEigenEnergies");
          e.upload(e.gamma, gammas[s], size_t(dims[s].NG) * np * np);
          e.have_vertex = true;
          e.upload(e.eps_i, eps_i[s], size_t(dims[s].No));
          e.upload(e.eps_a, eps_a[s], size_t(dims[s].Nv));
          e.have_eps = true;
          integrals_from_vertex(e);
          ccsd_begin(e, o, nullptr, nullptr);
          ccsd_iterate(e, o.max_iterations);
          energies[s] = e.last_energy;
          if (reports) fill_report(e, &reports[s]);
        });
        status[s] = rc;
        if (rc != SISI_OK) {
          msg[s] = g_last_error;
          energies[s] = std::nan("");
          if (reports) {
            reports[s] = sisi_ccsd_report{};
            reports[s].energy = std::nan("");
          }
        }
        if (reports) reports[s].status = rc;
      }
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    for (int s = 0; s < nsys; ++s)
      if (status[s] != SISI_OK)
        throw Error(status[s], "system " + std::to_string(s) + ": " + msg[s]);
  });
}

int sisi_engine_set_ladder_planes(sisi_engine *h, int enable) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    h->e.ladder_planes = enable != 0;
    h->e.drop_residual_graph();
  });
}

int sisi_ccsd_solve_lockstep(int nsys, const sisi_dims *dims, const double *const *gammas,
                             const double *const *eps_i, const double *const *eps_a,
                             const sisi_ccsd_opts *opts, double *energies,
                             sisi_ccsd_report *reports, sisi_lockstep_stats *stats) {
  return guarded([&] {
    SISI_REQUIRE(nsys >= 0 && dims && gammas && eps_i && eps_a && energies, SISI_ERR_INVALID,
                 "null argument");
    sisi_ccsd_opts o;
    if (opts) o = *opts;
    else sisi_ccsd_opts_default(&o);
    if (stats) *stats = sisi_lockstep_stats{};
    for (int s = 0; s < nsys; ++s) {
      SISI_REQUIRE(dims[s].No > 0 && dims[s].Nv > 0 && dims[s].NG > 0, SISI_ERR_INVALID,
                   "system " + std::to_string(s) + ": No, Nv, NG must be positive");
      SISI_REQUIRE(gammas[s] && eps_i[s] && eps_a[s], SISI_ERR_INVALID,
                   "system " + std::to_string(s) + ": null input");
      const size_t np = size_t(dims[s].No) + dims[s].Nv;
      try {
        check_finite(gammas[s], size_t(dims[s].NG) * np * np, "CoulombVertex");
        check_finite(eps_i[s], size_t(dims[s].No), "HoleEigenEnergies");
        check_finite(eps_a[s], size_t(dims[s].Nv), "ParticleEigenEnergies");
      } catch (const Error &err) {
        throw Error(err.code, "system " + std::to_string(s) + ": " + err.what());
      }
    }
    if (nsys == 0) return;
    ccsd_solve_lockstep(nsys, dims, gammas, eps_i, eps_a, o, energies, reports, stats);
  });
}

int sisi_engine_enable_stats(sisi_engine *h, int enable) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    h->e.timers->enabled = enable != 0;
  });
}

int sisi_engine_get_stats(sisi_engine *h, sisi_stats *stats) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    SISI_REQUIRE(stats, SISI_ERR_INVALID, "null pointer");
    h->e.timers->collect(*stats);
  });
}

int sisi_engine_reset_stats(sisi_engine *h) {
  return guarded([&] {
    ENGINE_REQUIRED(h);
    h->e.timers->reset();
  });
}

int sisi_generate_system(uint64_t seed, int No, int Nv, int NG, double *gamma, double *eps_i,
                         double *eps_a) {
  const int rc = sisi_generate_system_impl(seed, No, Nv, NG, gamma, eps_i, eps_a);
  g_last_error = rc == SISI_OK ? "" : "generate_system: No, Nv, NG must be positive";
  return rc;
}

int sisi_generate_ueg(int No, int Nv, double rs, int hartree_fock, double madelung, int *NG,
                      double *gamma, double *eps_i, double *eps_a) {
  const char *err = "";
  int rc;
  try {
    rc = sisi_generate_ueg_impl(No, Nv, rs, hartree_fock, madelung, NG, gamma, eps_i, eps_a, &err);
  } catch (const std::exception &ex) {
    g_last_error = std::string("generate_ueg: ") + ex.what();
    return SISI_ERR_ALLOC;
  }
  g_last_error = err;
  return rc;
}

int sisi_gemm_host(int precision, int transA, int transB, int M, int N, int K, double alpha,
                   const double *A, int lda, const double *B, int ldb, double beta, double *C,
                   int ldc) {
  return guarded([&] {
    SISI_REQUIRE(M >= 0 && N >= 0 && K >= 0, SISI_ERR_INVALID, "negative dimension");
    SISI_REQUIRE(is_precision(precision), SISI_ERR_INVALID, "unknown precision");
    SISI_REQUIRE(A && B && C, SISI_ERR_INVALID, "null pointer");
    const int64_t ra = transA ? K : M, ca = transA ? M : K;
    const int64_t rb = transB ? N : K, cb = transB ? K : N;
    SISI_REQUIRE(lda >= std::max<int64_t>(ra, 1) && ldb >= std::max<int64_t>(rb, 1) &&
                     ldc >= std::max(M, 1),
                 SISI_ERR_INVALID, "leading dimension too small");
    sisi_dims d{1, 1, 1};
    Engine e(d);
    // host elements actually present in a column-major (rows x cols, ld) matrix
    auto span = [](int64_t rows, int64_t cols, int64_t ld) -> int64_t {
      return (rows == 0 || cols == 0) ? 0 : ld * (cols - 1) + rows;
    };
    const int64_t na = span(ra, ca, lda), nb = span(rb, cb, ldb), nc = span(M, N, ldc);
    DBuf a(&e.pool, size_t(std::max<int64_t>(na, 1))), b(&e.pool, size_t(std::max<int64_t>(nb, 1))),
        c(&e.pool, size_t(std::max<int64_t>(nc, 1)));
    if (na > 0) e.upload(a, A, size_t(na));
    if (nb > 0) e.upload(b, B, size_t(nb));
    if (nc > 0) e.upload(c, C, size_t(nc));
    GemmArgs g;
    g.A = a.get();
    g.B = b.get();
    g.C = c.get();
    g.M = M;
    g.N = N;
    g.K = K;
    g.sam = transA ? lda : 1;
    g.sak = transA ? 1 : lda;
    g.sbk = transB ? ldb : 1;
    g.sbn = transB ? 1 : ldb;
    g.cm0 = 1;
    g.cn0 = ldc;
    g.dm0 = std::max(M, 1);
    g.dn0 = std::max(N, 1);
    g.alpha = alpha;
    g.beta = beta;
    gemm(e, precision, g, true);  // the precision asked for, whatever the size
    if (nc > 0) e.download(C, c, size_t(nc));
  });
}

int sisi_gemm_bench(int precision, int M, int N, int K, int reps, double *kernel_us,
                    double *total_us) {
  return guarded([&] {
    SISI_REQUIRE(M > 0 && N > 0 && K > 0 && reps > 0, SISI_ERR_INVALID, "bad bench shape");
    SISI_REQUIRE(is_precision(precision), SISI_ERR_INVALID, "unknown precision");
    sisi_dims d{1, 1, 1};
    Engine e(d);
    DBuf a(&e.pool, size_t(M) * K), b(&e.pool, size_t(K) * N), c(&e.pool, size_t(M) * N);
    fill_pattern(e, a.get(), int64_t(M) * K, 1);
    fill_pattern(e, b.get(), int64_t(K) * N, 2);
    GemmArgs g;
    g.A = a.get();
    g.B = b.get();
    g.C = c.get();
    g.M = M;
    g.N = N;
    g.K = K;
    g.sam = K;  // A(m,k) k-contiguous, B(k,n) k-contiguous: the planner's usual layouts
    g.sak = 1;
    g.sbk = 1;
    g.sbn = K;
    g.cm0 = 1;
    g.cn0 = M;
    g.dm0 = M;
    g.dn0 = N;
    gemm(e, precision, g, true);  // warm-up (and tuning, when enabled)
    e.sync();
    e.timers->enabled = true;
    e.timers->reset();
    hipEvent_t t0, t1;
    SISI_HIP(hipEventCreate(&t0));
    SISI_HIP(hipEventCreate(&t1));
    SISI_HIP(hipEventRecord(t0, e.stream));
    for (int r = 0; r < reps; ++r) gemm(e, precision, g, true);
    SISI_HIP(hipEventRecord(t1, e.stream));
    SISI_HIP(hipEventSynchronize(t1));
    float ms = 0.f;
    SISI_HIP(hipEventElapsedTime(&ms, t0, t1));
    SISI_HIP(hipEventDestroy(t0));
    SISI_HIP(hipEventDestroy(t1));
    sisi_stats st{};
    e.timers->collect(st);
    e.timers->enabled = false;
    const int kind = precision == SISI_FP64 ? 0 : precision == SISI_FP32 ? 1 : 2;
    if (kernel_us) *kernel_us = 1e6 * st.gemm_seconds[kind] / reps;
    if (total_us) *total_us = 1e3 * double(ms) / reps;
  });
}

}  // extern "C"
