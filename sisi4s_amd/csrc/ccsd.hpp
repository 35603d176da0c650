// CCSD on the device: integrals from the vertex, the real residual, the
// driver (update, mixers, energy, convergence).
#pragma once

#include "engine.hpp"

namespace sisi {

// CoulombIntegralsFromVertex::calculateRealIntegrals (cxx:203-571).
void integrals_from_vertex(Engine &e);

// CcsdEnergyFromCoulombIntegrals::getResiduum, real (cxx:27-494).  All
// pointers are device pointers; R1/R2 are overwritten.
void ccsd_residual(Engine &e, int iteration, bool initial_doubles, int slice_size,
                   const double *T1, const double *T2, double *R1, double *R2);

// Residual (FP64, or a bulk precision for the bulk phase's anchors) that
// records itself (and its intermediates) as the refinement anchor; and the
// residual at T by refinement around it.
void ccsd_residual_anchor(Engine &e, int slice_size, const double *T1, const double *T2,
                          double *R1, double *R2, int precision = SISI_FP64);
void ccsd_residual_delta(Engine &e, int slice_size, const double *T1, const double *T2, double *R1,
                         double *R2);

// getEnergy (ClusterSinglesDoublesAlgorithm.cxx:131-213).
void ccsd_energy(Engine &e, const double *T1, const double *T2, double &E, double &direct,
                 double &exchange, double *E_imag = nullptr);

// PairEnergy of getEnergy (ClusterSinglesDoublesAlgorithm.cxx:218-235,
// non-antisymmetrized): e_ij = sum_ab (T2 + T1 T1)_abij (2 V_abij - V_baij);
// pair is a device [No, No] buffer (i fastest; planar complex in complex mode).
void ccsd_pair_energies(Engine &e, const double *T1, const double *T2, double *pair);

// Complex path (complex mode, planar tensors):
// CoulombIntegralsFromVertex::calculateComplexIntegrals (cxx:613-810) and
// CcsdEnergyFromCoulombIntegrals::getResiduum for complex amplitudes
// (cxx:496-891); getEnergy and PairEnergy with F = complex
// (ClusterSinglesDoublesAlgorithm.cxx:131-235), E/direct/exchange as
// (Re, Im) and pair as a planar [No, No] buffer.
void zintegrals_from_vertex(Engine &e);
void zccsd_residual(Engine &e, int iteration, bool initial_doubles, int slice_size,
                    const double *T1, const double *T2, double *R1, double *R2);
void zccsd_energy(Engine &e, const double *T1, const double *T2, double E[2], double direct[2],
                  double exchange[2]);
void zccsd_pair_energies(Engine &e, const double *T1, const double *T2, double *pair);

// Coulomb factors (host arrays, interleaved complex128: lambda [NG, NR],
// pi [Np, NR]) for the factor-based ladder; NR = 0 removes them.  slice <= 0
// means Nv (factorsSliceSize default).
void set_coulomb_factors(Engine &e, int NR, int slice, const double *lambda, const double *pi,
                         const double *pi_out);
// R2 += the factor-based ladder of X = T2 + T1 T1 (real amplitudes, and
// planar complex amplitudes with the outgoing factor orbitals).
void factor_ladder(Engine &e, const double *T1, const double *X, double *R2, int precision);
void zfactor_ladder(Engine &e, const double *T1, const double *X, double *R2, int precision);

// Full-integral formulation (CcsdEnergyFromCoulombIntegralsReference.cxx:
// 29-295): integral classes given directly, PPPP/PPPH built from the vertex
// on request, and the residual from stored integrals.
void set_integrals(Engine &e, int which, const double *V);
// Direct-ring CCD (DrccdEnergyFromCoulombIntegrals.cxx:31-80)
void set_drccd(Engine &e, int unrestricted, int linearized, int adjacent_pairs_exchange);
void drccd_residual(Engine &e, int iteration, bool initial_doubles, const double *T2, double *R1,
                    double *R2);
// Integrals the current formulation needs are present.
bool integrals_ready(const Engine &e);
// Spin-orbital CCSD (UccsdAmplitudesFromCoulombIntegrals.cxx:107-241)
bool uccsd_ready(const Engine &e);
// the classes of the IP right-apply: those of uccsd_ready but PPPP
bool ip_eom_ready(const Engine &e);
// the classes of the EA right-apply: those of uccsd_ready but HHHH
bool ea_eom_ready(const Engine &e);
void uccsd_residual(Engine &e, int iteration, bool initial_doubles, const double *T1,
                    const double *T2, double *R1, double *R2);
double uccsd_energy(Engine &e, const double *T1, const double *T2);
// Amplitudes (host T1[a,i], T2[a,b,i,j]) of the similarity-transformed
// Hamiltonian of the EOM right-applies; the same amplitudes again keep the
// blocks already built, others drop them.
void eom_set_amplitudes(Engine &e, const double *T1, const double *T2);
// IP-EOM-CCSD right-apply (SimilarityTransformedHamiltonian.cxx:64-68): s =
// Hbar r on the 1h + 2h1p space, r1[i], r2[i,j,a] (device), spin-orbital
// engine; builds the IP blocks on first use
void eom_ip_sigma(Engine &e, const double *r1, const double *r2, double *s1, double *s2);
// EA-EOM-CCSD right-apply (SimilarityTransformedHamiltonian.cxx:289-294): s =
// Hbar r on the 1p + 2p1h space, r1[a], r2[j,a,b] (device), spin-orbital
// engine; builds the EA blocks on first use
void eom_ea_sigma(Engine &e, const double *r1, const double *r2, double *s1, double *s2);
void build_vabcd(Engine &e);
void ccsd_residual_full(Engine &e, int iteration, bool initial_doubles, const double *T1,
                        const double *T2, double *R1, double *R2);

// Driver (ClusterSinglesDoublesAlgorithm.cxx:38-128).
void ccsd_begin(Engine &e, const sisi_ccsd_opts &opts, const double *T1h, const double *T2h);
// Solution of the bordered DIIS system A x = b (DiisMixer.cxx:16-40) in FP64.
std::vector<double> diis_solve(std::vector<std::vector<double>> A, std::vector<double> b);
// Lockstep batched CCSD of nsys systems (inputs validated by the caller):
// every system's residual and update run in grouped launches shared by all
// systems.
void ccsd_solve_lockstep(int nsys, const sisi_dims *dims, const double *const *gammas,
                         const double *const *eps_i, const double *const *eps_a,
                         const sisi_ccsd_opts &o, double *energies, sisi_ccsd_report *reports,
                         sisi_lockstep_stats *stats);
void ccsd_iterate(Engine &e, int n);
void fill_report(const Engine &e, sisi_ccsd_report *r);

// Precision of the dense contractions: FP64 everywhere, FP32 everywhere, or
// FP32 with the particle-particle ladder in split BF16x3.
void set_precision(Engine &e, int precision);

}  // namespace sisi
