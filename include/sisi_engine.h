/*
 * sisi_engine.h -- C ABI of the MI355X CCSD / (T) contraction engine.
 *
 * Drop-in lower boundary for the CCSD / (T) hot path of sisi4s.  The
 * reference evaluates this path with Cyclops (CTF) expressions inside its
 * algorithm plugins; each entry point below replaces a named piece of that
 * path (cited as reference file:line).  The host-side plugin mirror
 * (CoulombIntegralsFromVertex, CcsdEnergyFromCoulombIntegrals,
 * CcsdPerturbativeTriples) calls only these functions.
 *
 * Conventions
 *  - Every function returns an int status: SISI_OK (0) or a SISI_ERR_* code.
 *    Nothing throws across the ABI; sisi_last_error() returns a thread-local
 *    message for the last failure on the calling thread.
 *  - All tensors are real FP64, dense, first index fastest (CTF global
 *    order): Gamma[G,p,q] at G + NG*(p + Np*q); T2[a,b,i,j] at
 *    a + Nv*(b + Nv*(i + No*j)); T1[a,i] at a + Nv*i.  p,q run over
 *    Np = No + Nv orbitals, holes first.
 *  - The caller owns host arrays; the engine owns device memory and one HIP
 *    stream.  A handle is not thread-safe (the reference drives the path
 *    from a single host thread per rank).
 *  - Complex mode: after sisi_engine_set_vertex_complex the engine runs the
 *    reference's complex path (getResiduum<complex>); every tensor argument
 *    of the CCSD entry points (amplitudes, residuals, integrals, pair
 *    energies) is then complex128, interleaved (re, im), with the same
 *    element order, i.e. twice as many doubles.  Eigenenergies stay real.
 *  - Precision: SISI_FP64 is the plain path.  SISI_FP32 and SISI_BF16X3 run
 *    the dense contractions on MFMA with FP32 accumulation; energies,
 *    denominators, the DIIS solve and all reductions stay FP64.
 */
#ifndef SISI_ENGINE_H
#define SISI_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SISI_OK 0
#define SISI_ERR_INVALID 1 /* bad argument / dimension */
#define SISI_ERR_HIP 2     /* HIP runtime error (no device, launch failure) */
#define SISI_ERR_STATE 3   /* call out of order (e.g. no vertex set) */
#define SISI_ERR_NUMERIC 4 /* NaN/Inf input or singular DIIS system */
#define SISI_ERR_ALLOC 5   /* device allocation failed */

#define SISI_FP64 0
#define SISI_FP32 1
#define SISI_BF16X3 2
#define SISI_BF16 3 /* one bf16 product (FP32 accumulation): refinement residuals */

#define SISI_MIXER_LINEAR 0
#define SISI_MIXER_DIIS 1

/* Integral classes (CoulombIntegralsFromVertex output names). */
#define SISI_PPHH 0 /* V_abij  [Nv,Nv,No,No] */
#define SISI_PHPH 1 /* V_aibj  [Nv,No,Nv,No] */
#define SISI_HHHH 2 /* V_ijkl  [No,No,No,No] */
#define SISI_HHHP 3 /* V_ijka  [No,No,No,Nv] */
#define SISI_HHPP 4 /* V_ijab  [No,No,Nv,Nv] */
#define SISI_PHHP 5 /* V_aijb  [Nv,No,No,Nv] (complex mode) */
#define SISI_PHHH 6 /* V_aijk  [Nv,No,No,No] (complex mode) */
#define SISI_PPPP 7 /* V_abcd  [Nv,Nv,Nv,Nv] (full-integral formulation) */
#define SISI_PPPH 8 /* V_abci  [Nv,Nv,Nv,No] (full-integral formulation) */

/* Residual formulations. */
#define SISI_FORMULATION_VERTEX 0    /* CcsdEnergyFromCoulombIntegrals       */
#define SISI_FORMULATION_INTEGRALS 1 /* CcsdEnergyFromCoulombIntegralsReference */
#define SISI_FORMULATION_ONLY_PPL 2  /* ... its onlyPPL branch                */
#define SISI_FORMULATION_DRCCD 3     /* DrccdEnergyFromCoulombIntegrals       */
#define SISI_FORMULATION_UCCSD 4     /* UccsdAmplitudesFromCoulombIntegrals   */

typedef struct sisi_engine sisi_engine;

typedef struct {
  int No; /* occupied orbitals   (HoleEigenEnergies length)     */
  int Nv; /* virtual orbitals    (ParticleEigenEnergies length) */
  int NG; /* auxiliary index     (CoulombVertex first length)   */
} sisi_dims;

/* Arguments of the CCSD driver; names and defaults follow
 * ClusterSinglesDoublesAlgorithm.hpp:42-46, .cxx:48-83, DiisMixer.cxx:55,
 * LinearMixer.cxx:17, CcsdEnergyFromCoulombIntegrals.cxx:374-384. */
typedef struct {
  int mixer;                     /* SISI_MIXER_LINEAR (default) or _DIIS  */
  int max_residua;               /* maxResidua, default 4                 */
  double mixing_ratio;           /* mixingRatio, default 1.0              */
  int max_iterations;            /* maxIterations, default 16             */
  double amplitudes_convergence; /* amplitudesConvergence, default 1e-5   */
  double energy_convergence;     /* energyConvergence, default 1e-6       */
  double level_shift;            /* levelShift, default 0                 */
  int slice_size;                /* integralsSliceSize, <= 0 means No     */
  int precision;                 /* bulk precision: SISI_FP64/_FP32/_BF16X3 */
  double switch_threshold;       /* |dT|/|T| at which a mixed run switches
                                    to FP64 iterations (default 1e-5)      */
  int refine_interval;           /* after the switch: 0 = plain FP64
                                    iterations; k > 0 = an FP64 anchor
                                    residual every k iterations and
                                    refinement residuals (anchor + reduced-
                                    precision change) in between          */
  int refine_precision;          /* precision of those refinement residuals:
                                    SISI_BF16 (default), SISI_BF16X3 or
                                    SISI_FP32; their error scales with the
                                    change since the anchor               */
  int ppl;                       /* ppl (CcsdEnergyFromCoulombIntegrals.cxx:
                                    324-338), default 1; 0 drops the
                                    particle-particle ladder              */
  int distinguishable;           /* distinguishable (cxx:123-124), default
                                    0; nonzero selects the distinguishable-
                                    cluster terms (cxx:150-187, 244-246,
                                    283-285, 316-320)                     */
  double coarse_threshold;       /* mixed runs: a coarse stage in
                                    coarse_precision runs before the bulk
                                    precision until |dT|/|T| <= this value
                                    (default 0: no coarse stage)          */
  int coarse_precision;          /* SISI_BF16 (default), SISI_FP32 or
                                    SISI_BF16X3                            */
  int graphs;                    /* nonzero: the plain residual is captured
                                    once per precision in a hipGraph and
                                    replayed (small, launch-bound systems);
                                    default 0                             */
  int bulk_refine_interval;      /* before the switch: 0 = plain bulk-
                                    precision iterations (default); k > 0 =
                                    an anchor residual in the bulk precision
                                    every k iterations and refinement
                                    residuals in refine_precision between
                                    (vertex formulation, real amplitudes) */
} sisi_ccsd_opts;

typedef struct {
  double energy;      /* correlation energy after the last iteration     */
  double direct;      /* direct term of getEnergy                        */
  double exchange;    /* exchange term of getEnergy                      */
  double delta;       /* |dT|/|T| of the last iteration                  */
  int iterations;     /* iterations run so far                           */
  int converged;      /* reference convergence test met                  */
  int fp64_iterations;/* iterations run with FP64 contractions           */
  int precision;      /* precision of the last iteration                 */
  int status;         /* SISI_OK, or the error of this system in a batch */
  double energy_imag; /* imaginary part of the energy (complex mode; the
                         reference reports the real part, cxx:30-33)     */
} sisi_ccsd_report;

typedef struct {
  double gemm_seconds[3];   /* device time of MFMA GEMM launches per precision */
  double gemm_flops[3];     /* algorithmic flops of those launches             */
  int64_t gemm_launches[3]; /* launch count per precision                      */
  double other_seconds;     /* device time of the bandwidth kernels           */
  double other_bytes;       /* algorithmic bytes of the bandwidth kernels     */
  int64_t other_launches;
} sisi_stats;

/* Last error message of the calling thread ("" if none). */
const char *sisi_last_error(void);

/* Fill opts with the reference defaults (see sisi_ccsd_opts). */
void sisi_ccsd_opts_default(sisi_ccsd_opts *opts);

/* Create an engine on the current HIP device.  Replaces the implicit CTF
 * world of the reference run loop (src/Sisi4s.cxx:280-282). */
int sisi_engine_create(const sisi_dims *dims, sisi_engine **out);
int sisi_engine_destroy(sisi_engine *e);

/* Upload the real Coulomb vertex Gamma[NG,Np,Np] and the eigenenergies.
 * Replaces the reads of "CoulombVertex", "HoleEigenEnergies" and
 * "ParticleEigenEnergies" (CoulombIntegralsFromVertex.cxx:23-38,
 * CcsdEnergyFromCoulombIntegrals.cxx:58-115). */
int sisi_engine_set_vertex(sisi_engine *e, const double *gamma);
int sisi_engine_set_eigenenergies(sisi_engine *e, const double *eps_i,
                                  const double *eps_a);

/* Upload a complex vertex Gamma[NG,Np,Np] (complex128, interleaved) and
 * switch the engine to complex mode: CoulombIntegralsFromVertex with
 * complex: 1 (calculateComplexIntegrals, CoulombIntegralsFromVertex.cxx:
 * 613-810) and the complex residual (CcsdEnergyFromCoulombIntegrals.cxx:
 * 496-891, ClusterSinglesDoublesAlgorithm.cxx:490-565), which the
 * reference selects when the integrals are complex (cxx:25-35).  No
 * symmetry of Gamma is assumed.  sisi_engine_set_vertex switches back. */
int sisi_engine_set_vertex_complex(sisi_engine *e, const double *gamma);

/* CoulombIntegralsFromVertex::calculateRealIntegrals
 * (CoulombIntegralsFromVertex.cxx:203-571): builds PPHH, PHPH, HHHH, HHHP
 * (and HHPP as a transposed view) on the device. */
int sisi_integrals_from_vertex(sisi_engine *e);

/* Copy one integral class (SISI_PPHH ...) to host memory (complex128 in
 * complex mode, where HHPP, PHHP and PHHH are the reference's own classes).
 * PPPP and PPPH (CoulombIntegralsFromVertex.cxx:423, 430) are built from the
 * real vertex on request. */
int sisi_get_integrals(sisi_engine *e, int which, double *out);

/* Precision used by the dense contractions of the next residual call. */
int sisi_engine_set_precision(sisi_engine *e, int precision);

/* CoulombFactors Lambda[G,R] (complex128 [NG, NR]), FactorOrbitals
 * Pi[p,R] and OutgoingFactorOrbitals Po[p,R] (complex128 [Np, NR]; NULL
 * means Pi): while set (NR > 0), the particle-particle ladder is the
 * factor-based one, sliceAmplitudesFromCoupledCoulombFactors
 * (ClusterSinglesDoublesAlgorithm.cxx:567-672 for real amplitudes, which
 * reads Pi only; 745-880 for complex amplitudes) over factor slices of
 * factors_slice_size (<= 0: Nv, the factorsSliceSize default,
 * CcsdEnergyFromCoulombIntegrals.cxx:340-372, 786-813).  It equals the
 * vertex ladder when Gamma[G,q,r] = sum_R Lambda_GR Po_qR Pi_rR.  NR = 0
 * removes the factors.  Not combined with the refinement residuals. */
int sisi_engine_set_coulomb_factors(sisi_engine *e, int NR, int factors_slice_size,
                                    const double *lambda, const double *pi,
                                    const double *pi_out);

/* Upload one real integral class (SISI_PPHH, _PHPH, _HHHH, _HHHP, _PPPP,
 * _PPPH; _PHHP and _HHPP for the direct-ring CCD formulation) instead of
 * building it from a vertex: the inputs of
 * CcsdEnergyFromCoulombIntegralsReference, which reads integrals only.
 * With PPHH, PHPH, HHHH and HHHP given the engine counts as having its
 * integrals (no vertex needed for the full-integral formulation). */
int sisi_engine_set_integrals(sisi_engine *e, int which, const double *V);

/* Residual formulation of the next residual calls and solves:
 * SISI_FORMULATION_VERTEX (default), SISI_FORMULATION_INTEGRALS
 * (CcsdEnergyFromCoulombIntegralsReference::getResiduum, cxx:29-295, stored
 * V_abcd and V_abci) or SISI_FORMULATION_ONLY_PPL (its onlyPPL branch,
 * cxx:56-75: R2 = V_abcd (T2 + T1 T1), R1 = 0, every iteration) or
 * SISI_FORMULATION_DRCCD (DrccdEnergyFromCoulombIntegrals::getResiduum,
 * DrccdEnergyFromCoulombIntegrals.cxx:31-80: doubles only, from PPHH, PHHP
 * and HHPP; needs only PPHH uploaded or a vertex) or SISI_FORMULATION_UCCSD
 * (UccsdAmplitudesFromCoulombIntegrals::getResiduumTemplate,
 * UccsdAmplitudesFromCoulombIntegrals.cxx:107-241: spin-orbital CCSD on an
 * engine whose No, Nv count spin orbitals, from the antisymmetrized classes
 * <pq||rs> PPHH, HHHH, PPPP, HHHP, PPPH and PHPH uploaded with
 * sisi_engine_set_integrals; energy 1/4 T2 V + 1/2 T1 T1 V, reported as
 * `direct`, exchange 0). */
int sisi_engine_set_formulation(sisi_engine *e, int formulation);

/* CCSD amplitudes T1[Nv,No], T2[Nv,Nv,No,No] of the EOM right-applies below
 * (SimilarityTransformedHamiltonian's amplitudes, cxx:64-68): the
 * similarity-transformed Hamiltonian's blocks are built from them once, on
 * the first IP / EA right-apply, and kept on the engine until other
 * amplitudes, integrals or eigenenergies are set. */
int sisi_eom_set_amplitudes(sisi_engine *e, const double *T1, const double *T2);

/* IP-EOM-CCSD right-apply, the product UCcsdIPEquationOfMotionDavidson
 * (UCcsdIPEquationOfMotionDavidson.cxx:43-205) repeats in its Davidson
 * iterations: SimilarityTransformedHamiltonian::right_apply_CCSD_IP
 * (SimilarityTransformedHamiltonian.cxx:64-68) with CCSD dressing, FP64, on an
 * engine in SISI_FORMULATION_UCCSD (spin orbitals, canonical orbitals).
 * T1[Nv,No], T2[Nv,Nv,No,No]: the CCSD amplitudes (as sisi_eom_set_amplitudes;
 * both NULL: the amplitudes set before); r1[No], r2[No,No,Nv] (antisymmetric
 * in the two holes), first index fastest; s1, s2 alike. */
int sisi_eom_ip_sigma(sisi_engine *e, const double *T1, const double *T2, const double *r1,
                      const double *r2, double *s1, double *s2);

/* EA-EOM-CCSD right-apply, the product UCcsdEAEquationOfMotionDavidson
 * (UCcsdEAEquationOfMotionDavidson.cxx:43-351) repeats in its Davidson
 * iterations: SimilarityTransformedHamiltonian::right_apply_CCSD_EA
 * (SimilarityTransformedHamiltonian.cxx:289-294), FP64, on an engine in
 * SISI_FORMULATION_UCCSD with PPHH, PPPP, HHHP, PPPH and PHPH uploaded.
 * T1, T2 as for sisi_eom_ip_sigma; r1[Nv], r2[No,Nv,Nv] (antisymmetric in the
 * two particles), first index fastest; s1, s2 alike. */
int sisi_eom_ea_sigma(sisi_engine *e, const double *T1, const double *T2, const double *r1,
                      const double *r2, double *s1, double *s2);

/* Arguments of DrccdEnergyFromCoulombIntegrals (cxx:41-65): unrestricted
 * (spins = 1 instead of 2, also in the energy of
 * ClusterSinglesDoublesAlgorithm.cxx:133,163-170), linearized (no quadratic
 * term) and adjacentPairsExchange (W_ijab = spins V_ijab - V_jiab).  Applies
 * to SISI_FORMULATION_DRCCD; defaults 0, 0, 0.  Replaces the integer
 * arguments read in DrccdEnergyFromCoulombIntegrals.cxx:41,47,63. */
int sisi_engine_set_drccd(sisi_engine *e, int unrestricted, int linearized,
                          int adjacent_pairs_exchange);

/* Energy variants of getEnergy (ClusterSinglesDoublesAlgorithm.cxx:131-213)
 * for the next energies of this engine (driver, sisi_ccsd_energy, pair
 * energies): unrestricted = 1 sets spins = 1 (cxx:133) in the direct and
 * exchange terms; antisymmetrize = 1 evaluates 0.25 T_abkl V_klab +
 * 0.5 T_aj T_cl V_jlac (cxx:163-167) and its pair energies (cxx:222-226),
 * real amplitudes only.  Replaces the `unrestricted` and `antisymmetrize`
 * arguments of the CCSD plugins. */
int sisi_engine_set_energy_variant(sisi_engine *e, int unrestricted, int antisymmetrize);

/* HPFockMatrix f[i,a] ([No, Nv], real vertex; NULL clears):
 * every energy adds the non-canonical term spins * sum_ai T_ai f_ia
 * (ClusterSinglesDoublesAlgorithm.cxx:199-211); its last value is returned
 * by sisi_engine_get_noncanonical_energy. */
int sisi_engine_set_hp_fock(sisi_engine *e, const double *fia);
int sisi_engine_get_noncanonical_energy(sisi_engine *e, double *out);

/* Residual variant of the next residual calls: the ppl and distinguishable
 * arguments of CcsdEnergyFromCoulombIntegrals (cxx:123-124, 324-338);
 * default ppl = 1, distinguishable = 0.  sisi_ccsd_begin / _solve set it
 * from their opts. */
int sisi_engine_set_residual_variant(sisi_engine *e, int ppl, int distinguishable);

/* CcsdEnergyFromCoulombIntegrals::getResiduum, real path
 * (CcsdEnergyFromCoulombIntegrals.cxx:27-494) with the vertex-built ladder
 * (cxx:336-453, ClusterSinglesDoublesAlgorithm.cxx:887-915).
 * iteration == 0 and !initial_doubles gives the MP2 branch (cxx:47-51). */
int sisi_ccsd_residual(sisi_engine *e, int iteration, int initial_doubles,
                       int slice_size, const double *T1, const double *T2,
                       double *R1, double *R2);

/* The residual of sisi_ccsd_residual (iteration >= 1) in two parts used by
 * the mixed-precision refinement: _anchor evaluates it in FP64 at T and
 * records T, R and the T-dependent intermediates as the anchor; _refine
 * evaluates it at a new T as R(T_anchor) + [R(T) - R(T_anchor)], with the
 * bracket expanded so every product carries a factor T - T_anchor and
 * computed in the precision of sisi_engine_set_precision.  No counterpart
 * in the reference (its residual is always one FP64 evaluation). */
int sisi_ccsd_residual_anchor(sisi_engine *e, int slice_size, const double *T1,
                              const double *T2, double *R1, double *R2);
int sisi_ccsd_residual_refine(sisi_engine *e, int slice_size, const double *T1,
                              const double *T2, double *R1, double *R2);

/* ClusterSinglesDoublesAlgorithm::getEnergy (cxx:131-213), restricted,
 * non-antisymmetrized: out[0] = E, out[1] = direct, out[2] = exchange. */
int sisi_ccsd_energy(sisi_engine *e, const double *T1, const double *T2,
                     double out[3]);

/* MP2 energy as the reference obtains it: iteration 0 of the driver
 * (R2 = V_abij, T2 = -R2/D) followed by getEnergy. */
int sisi_mp2_energy(sisi_engine *e, double *energy);

/* ClusterSinglesDoublesAlgorithm::run<double> (cxx:38-128), split so a
 * caller can time iterations: begin() sets up amplitudes and the mixer
 * (createAmplitudes cxx:240-273; T1/T2 NULL means zero amplitudes, i.e. the
 * MP2 start), iterate() runs up to n iterations and stops early when the
 * reference convergence test (cxx:103-106) is met. */
int sisi_ccsd_begin(sisi_engine *e, const sisi_ccsd_opts *opts,
                    const double *T1_init, const double *T2_init);
int sisi_ccsd_iterate(sisi_engine *e, int n, sisi_ccsd_report *report);

/* begin + iterate(max_iterations): the whole CcsdEnergyFromCoulombIntegrals
 * run (ClusterSinglesDoublesAlgorithm.cxx:25-35). */
int sisi_ccsd_solve(sisi_engine *e, const sisi_ccsd_opts *opts,
                    const double *T1_init, const double *T2_init,
                    sisi_ccsd_report *report);

/* storeAmplitudes (cxx:288-299): copy the current T1/T2 to the host. */
int sisi_ccsd_get_amplitudes(sisi_engine *e, double *T1, double *T2);

/* PairEnergy of getEnergy (ClusterSinglesDoublesAlgorithm.cxx:218-235,
 * non-antisymmetrized branch) for the current amplitudes:
 * pair[i + No*j] = sum_ab tau_abij (2 V_abij - V_baij), tau = T2 + T1 T1;
 * the pairs sum to the correlation energy. */
int sisi_ccsd_pair_energies(sisi_engine *e, double *pair);

/* Per-iteration energies of the current run (energy.value per iteration,
 * cxx:185-213); *n receives the count (at most max). */
int sisi_ccsd_get_energies(sisi_engine *e, double *energies, int max, int *n);

/* CcsdPerturbativeTriples::run (CcsdPerturbativeTriples.cxx:119-248):
 * E_(T) (without the CCSD energy) for amplitudes T1/T2 (NULL: the engine's
 * current amplitudes).  precision selects the triples contractions;
 * energies are always accumulated in FP64. */
int sisi_triples_energy(sisi_engine *e, const double *T1, const double *T2,
                        int precision, double *energy);

/* Same, restricted to the nblocks sorted blocks (i0<=i1<=i2) listed in
 * blocks[3*b..3*b+2]; used to time a sample of the triple loop. */
int sisi_triples_energy_blocks(sisi_engine *e, const double *T1,
                               const double *T2, int precision,
                               const int *blocks, int nblocks, double *energy);

/* Many independent CCSD runs (batches of small systems of uneven
 * No/Nv/NG) on one GPU: systems are taken from a shared queue by
 * `workers` host threads, each driving its own engine and HIP stream, so the
 * small kernels of different systems overlap on the device.  Each system is
 * ClusterSinglesDoublesAlgorithm::run (cxx:38-128) with the same opts.
 * dims/gammas/eps_i/eps_a hold nsys entries; energies[s] and reports[s]
 * (may be NULL) receive the results (a failed system gets NaN and its status
 * in reports[s].status).  Returns the first failing status. */
int sisi_ccsd_solve_batched(int nsys, const sisi_dims *dims, const double *const *gammas,
                            const double *const *eps_i, const double *const *eps_a,
                            const sisi_ccsd_opts *opts, int workers, double *energies,
                            sisi_ccsd_report *reports);

/* Reduced-precision ladder with plane output (default on; no counterpart
 * in the reference): the split-BF16x3 / bf16 V~_xycd slice-pair product
 * (CcsdEnergyFromCoulombIntegrals.cxx:431-434) writes the bf16 hi/lo planes
 * the X contraction (cxx:443) reads, with the rounding an operand split of
 * the FP64 slice pair would give; off: the FP64 slice pair is stored and
 * split by the next product.  Results are the same either way. */
int sisi_engine_set_ladder_planes(sisi_engine *e, int enable);

/* Counters of a lockstep batched run (sisi_ccsd_solve_lockstep). */
typedef struct {
  double iterations; /* lockstep iterations (the longest system's count)  */
  double launches;   /* grouped kernel launches, all iterations            */
  double records;    /* per-system launch records those launches covered   */
  double plans;      /* residual schedules built (precision switches and
                        systems leaving the batch rebuild it)             */
  double setup_seconds;    /* engines, integrals, recorded residuals (wall) */
  double residual_seconds; /* device time of the grouped residuals          */
  double update_seconds;   /* device time of the grouped updates            */
  double total_seconds;    /* the whole call (wall)                          */
} sisi_lockstep_stats;

/* The batch of sisi_ccsd_solve_batched iterated in lockstep instead of on
 * host threads (no counterpart in the reference, which runs one system per
 * driver invocation, ClusterSinglesDoublesAlgorithm.cxx:38-128): each
 * system's residual is recorded once per precision as a list of launches,
 * and every launch of an iteration (residual, amplitude estimate, mixer
 * update, norms, energy) is one grouped launch over all systems, replayed
 * from a hipGraph while the set of systems and their precisions stay the
 * same.  The number of launches per iteration does not grow with nsys.
 * Per system the iteration is ClusterSinglesDoublesAlgorithm::run with the
 * same opts; reduced precisions (bulk and coarse stages of a mixed run)
 * compute their products on FP32 MFMA, and the report's precision says so.
 * Requires refine_interval == 0 and max_residua <= 30.  stats may be NULL. */
int sisi_ccsd_solve_lockstep(int nsys, const sisi_dims *dims, const double *const *gammas,
                             const double *const *eps_i, const double *const *eps_a,
                             const sisi_ccsd_opts *opts, double *energies,
                             sisi_ccsd_report *reports, sisi_lockstep_stats *stats);

/* Device timing of kernel families since the last reset (collected only
 * while enabled; enabling adds one event pair per launch). */
int sisi_engine_enable_stats(sisi_engine *e, int enable);
int sisi_engine_get_stats(sisi_engine *e, sisi_stats *stats);
int sisi_engine_reset_stats(sisi_engine *e);

/* Synthetic seeded system (host, no device needed): Philox-4x64-10 vertex
 * Gamma[G,p,q] = Gamma[G,q,p] ~ N(0,1) exp(-G/100)/sqrt(NG), eps_i in
 * U(-2,-0.5), eps_a in U(0.2,3).  Replaces the reference's random-tensor
 * generators (GenerateRandomTensor.cxx:19-29, math/RandomTensor.hpp:38-58),
 * whose mt19937-per-rank streams are not reproducible. */
int sisi_generate_system(uint64_t seed, int No, int Nv, int NG, double *gamma,
                         double *eps_i, double *eps_a);

/* Uniform electron gas (host, no device needed): UegVertexGenerator
 * (UegVertexGenerator.cxx:49-229; arguments No, Nv, rs, hartreeFock,
 * madelung < 0 to evaluate it) with the plane-wave vertex rewritten in real
 * orbitals, cos/sin pairs of (k, -k), and real auxiliary functions, cos/sin
 * pairs of (G, -G), so that Gamma[G,p,q] == Gamma[G,q,p] as the real path
 * requires; correlation energies equal the plane-wave ones.  Needs closed
 * shells at No and No+Nv.  Call with gamma == NULL to get *NG; then pass
 * gamma of NG*Np*Np doubles with *NG holding that count. */
int sisi_generate_ueg(int No, int Nv, double rs, int hartree_fock, double madelung, int *NG,
                      double *gamma, double *eps_i, double *eps_a);

/* Backend of the split-BF16x3 and bf16 products of one operand pair (batch
 * 1) for the whole process: 0 = the engine's own MFMA kernel (default), 1 =
 * plain bf16 GEMMs of hipBLASLt on the engine's operand planes (K tripled,
 * FP32 accumulation) followed by the engine's FP64 epilogue.  The
 * environment variable SISI_B3_BACKEND=library sets 1 at start-up.  No
 * counterpart in the reference. */
int sisi_set_bf16_backend(int library);

/* Kernel variant of the split-BF16x3 (passes = 3) or bf16 (passes = 1)
 * products for the whole process: 0 .. sisi_gemm_variant_count(passes)-1
 * forces one, -1 returns to the automatic choice (tuning cache, else a size
 * rule).  The environment variables SISI_B3_VARIANT / SISI_B1_VARIANT set it
 * at start-up.  No counterpart in the reference (CTF picks BLAS kernels). */
int sisi_gemm_set_variant(int passes, int variant);
int sisi_gemm_variant_count(int passes);
const char *sisi_gemm_variant_name(int passes, int variant);

/* Autotuner of the GEMM kernel variants: with tuning on (or SISI_TUNE=1),
 * the first product of each (family, log2 M, log2 N, log2 K) class times
 * every variant on the device and records the fastest in the process and
 * in the cache file (SISI_TUNE_CACHE, else tuning/gfx950.txt beside the
 * library); recorded classes are never timed again.  sisi_tune_dump writes
 * the records as text (needed = bytes including the terminating 0).  No
 * counterpart in the reference. */
int sisi_tune_enable(int on);
int sisi_tune_dump(char *buf, int len, int *needed);

/* C = alpha*op(A)*op(B) + beta*C (column-major, FP64 or FP32 MFMA) on
 * device arrays the caller owns, on the caller's HIP stream (NULL: the null
 * stream), with no allocation of any kind: the split-K partial sums live in
 * a 32 MiB device array inside the library (one per GPU, held until the
 * process ends) whose address is looked up on the first call on each
 * device.  ksplit = 0 picks the split sisi_gemm_host would (fewer partial
 * sums when they would not fit); ksplit > 0 asks for that split and fails
 * with SISI_ERR_INVALID when ksplit*M*N doubles exceed the array.  Calls
 * that overlap on one GPU share the array and race.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, allocation-free harnesses). */
int sisi_gemm_device_noalloc(int precision, int transA, int transB, int M, int N, int K,
                             double alpha, const double *A, int lda, const double *B,
                             int ldb, double beta, double *C, int ldc, int ksplit,
                             void *stream);

/* Timing of one GEMM shape on device-resident operands (A m x k and B k x n,
 * both k-contiguous, C column-major): average microseconds per call of the
 * GEMM kernels (HIP events around each launch) and of the whole call
 * (operand splits included), over `reps` calls after one warm-up.  No
 * counterpart in the reference. */
int sisi_gemm_bench(int precision, int M, int N, int K, int reps, double *kernel_us,
                    double *total_us);

/* One MFMA GEMM C = alpha*op(A)*op(B) + beta*C on host arrays (column-major,
 * op = transpose when trans* != 0), in the given precision.  No counterpart
 * in the reference (CTF calls BLAS internally); exposed so the kernel family
 * every contraction uses can be tested on its own. */
int sisi_gemm_host(int precision, int transA, int transB, int M, int N, int K,
                   double alpha, const double *A, int lda, const double *B,
                   int ldb, double beta, double *C, int ldc);

#ifdef __cplusplus
}
#endif

#endif /* SISI_ENGINE_H */
