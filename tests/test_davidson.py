"""The host Davidson driver of the EOM eigenproblems (sisi4s_amd/host/davidson.hpp,
the role of EigenSystemDavidsonMono, math/EigenSystemDavidson.hpp:114-428) on the CPU:
the small non-symmetric eigensolver against closed forms and numpy's dgeev, and
the Davidson iteration against the full spectrum of dense matrices, including
the IP-EOM-CCSD matrix of the oracle, whose two-electron roots equal full CI."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "davidson_check.cpp")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path_factory.mktemp("davidson") / "davidson_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", SRC, "-o", exe], check=True)
    return exe


def _run(exe, text):
    r = subprocess.run([exe], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def eigenvalues(exe, A):
    n = A.shape[0]
    text = "eig %d\n%s\n" % (n, " ".join("%.17g" % x for x in A.ravel()))
    return np.array([complex(*map(float, l.split())) for l in _run(exe, text)])


def davidson(exe, A, diag, nroots, tol_vectors=1e-20, tol_energy=1e-13, max_basis=1000,
             max_iterations=100, min_iterations=1, refresh=0):
    n = A.shape[0]
    text = "davidson %d %d %.17g %.17g %d %d %d %d\n%s\n%s\n" % (
        n, nroots, tol_vectors, tol_energy, max_basis, max_iterations, min_iterations, refresh,
        " ".join("%.17g" % x for x in A.ravel()), " ".join("%.17g" % x for x in diag))
    lines = _run(exe, text)
    vals = np.array([complex(*map(float, l.split())) for l in lines[:nroots]])
    resid = np.array([float(l.split()[1]) for l in lines[nroots:2 * nroots]])
    tail = lines[-1].split()
    info = dict(iterations=int(tail[1]), converged=bool(int(tail[3])), rms=float(tail[5]),
                applies=int(tail[7]))
    return vals, resid, info


def _match(a, b):
    """Largest distance from each value of b to its nearest value in a."""
    return max(np.min(np.abs(a - z)) for z in b)


def test_small_eigenvalues_closed_forms(harness):
    # companion matrix of prod (x - r_k): eigenvalues are the roots
    roots = np.array([-3.0, -1.5, 0.25, 1.0, 2.0, 4.5])
    c = np.poly(roots)
    C = np.zeros((6, 6))
    C[0, :] = -c[1:]
    C[1:, :-1] = np.eye(5)
    ev = eigenvalues(harness, C)
    assert np.abs(ev.imag).max() <= 1e-9
    assert np.abs(np.sort(ev.real) - roots).max() <= 1e-9
    # rotation by theta embedded with two real eigenvalues: e^{+-i theta}, 2, -1
    th = 0.7
    R = np.zeros((4, 4))
    R[:2, :2] = [[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]
    R[2, 2], R[3, 3], R[0, 3], R[2, 1] = 2.0, -1.0, 0.3, -0.4
    Q, _ = np.linalg.qr(np.random.default_rng(3).normal(size=(4, 4)))
    ev = eigenvalues(harness, Q @ R @ Q.T)
    want = np.array([np.exp(1j * th), np.exp(-1j * th), 2.0, -1.0])
    assert _match(ev, want) <= 1e-10 and _match(want, ev) <= 1e-10
    # upper triangular: the diagonal
    T = np.triu(np.random.default_rng(4).normal(size=(7, 7)))
    assert _match(eigenvalues(harness, T), np.diag(T)) <= 1e-10


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (5, 2), (17, 3), (40, 4), (90, 5)])
def test_small_eigenvalues_against_dgeev(harness, n, seed):
    A = np.random.default_rng(seed).normal(size=(n, n))
    ev = eigenvalues(harness, A)
    ref = np.linalg.eigvals(A)
    assert len(ev) == n
    scale = max(1.0, np.abs(ref).max())
    assert _match(ev, ref) <= 1e-9 * scale and _match(ref, ev) <= 1e-9 * scale
    assert abs(ev.sum() - np.trace(A)) <= 1e-9 * scale * n


def _diagonally_dominant(n, seed):
    """Non-symmetric, well-separated diagonal plus weak couplings: the lowest
    eigenvalues are real."""
    rng = np.random.default_rng(seed)
    d = np.linspace(-2.0, 6.0, n) + 0.005 * rng.uniform(size=n)
    A = np.diag(d) + 0.004 * rng.normal(size=(n, n))
    ref = np.sort_complex(np.linalg.eigvals(A))
    assert np.abs(ref[:6].imag).max() == 0.0
    return A, d, ref.real


def test_davidson_lowest_roots_of_a_nonsymmetric_matrix(harness):
    A, d, ref = _diagonally_dominant(300, 11)
    k = 4
    vals, resid, info = davidson(harness, A, d, k, tol_vectors=1e-20, tol_energy=0.0)
    assert info["converged"] and info["iterations"] < 60
    assert np.abs(vals.real - ref[:k]).max() <= 1e-11
    assert resid.max() <= 1e-9
    # one right-apply per basis vector: far fewer than the reference's
    # basis-times-iterations products
    assert info["applies"] <= k * (info["iterations"] + 1)
    # the reference's stopping rule: either criterion ends the iteration
    vals, _, info = davidson(harness, A, d, k, tol_vectors=1e-30, tol_energy=1e-3)
    assert info["converged"] and info["iterations"] <= 3


def test_davidson_stops_at_the_iteration_limit_and_basis_size(harness):
    A, d, ref = _diagonally_dominant(120, 12)
    _, _, info = davidson(harness, A, d, 2, tol_vectors=1e-30, tol_energy=0.0, max_iterations=3)
    assert info["iterations"] == 3 and not info["converged"]
    _, _, info = davidson(harness, A, d, 2, tol_vectors=1e-30, tol_energy=0.0, max_basis=8)
    assert not info["converged"] and info["iterations"] <= 4  # basis 2, 4, 6, 8, 10
    # with refresh the basis restarts from the Ritz vectors and converges
    vals, _, info = davidson(harness, A, d, 2, tol_vectors=1e-20, tol_energy=0.0, max_basis=8,
                             max_iterations=400, refresh=1)
    assert info["converged"]
    assert np.abs(vals.real - ref[:2]).max() <= 1e-10


def test_davidson_min_iterations(harness):
    n = 30
    d = np.arange(n, dtype=float)
    A = np.diag(d)
    vals, _, info = davidson(harness, A, d, 1, min_iterations=5)
    assert info["iterations"] == 5 and abs(vals[0]) <= 1e-14


def _ip_problem(no_spatial, nv_spatial, seed, scale):
    from oracle import eom, philox, uccsd
    g, ei, ea = philox.system(seed, no_spatial, nv_spatial, 12)
    g = scale * g
    ints = uccsd.spin_orbital_integrals(g, no_spatial, nv_spatial)
    epi, epa = np.repeat(ei, 2), np.repeat(ea, 2)
    r = uccsd.solve(ints, epi, epa, max_residua=6, max_iterations=200,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert r["converged"]
    H = eom.hbar_ip(ints, epi, epa, r["T1"], r["T2"])
    no, nv = len(epi), len(epa)
    M = eom.ip_matrix(H, no, nv)
    # IPCcsdPreconditioner's diagonal (CcsdPreconditioner.cxx:546-560)
    pairs = [(i, j) for i in range(no) for j in range(i + 1, no)]
    diag = np.concatenate([-epi] + [-epi[i] - epi[j] + epa for (i, j) in pairs])
    return M, diag


def test_davidson_ip_eom_closed_shell(harness):
    """Lowest IP-EOM-CCSD roots of a closed-shell system (No=3, Nv=5 spatial,
    dimension 6 + 15 x 10) by Davidson against the dense spectrum: the
    spin doublets come out degenerate."""
    M, diag = _ip_problem(3, 5, 61, 0.3)
    ref = np.sort(np.linalg.eigvals(M).real)
    vals, resid, info = davidson(harness, M, diag, 4, tol_vectors=1e-20, tol_energy=0.0)
    assert info["converged"]
    assert np.abs(vals.real - ref[:4]).max() <= 1e-9
    assert abs(vals[0].real - vals[1].real) <= 1e-9 and abs(vals[2].real - vals[3].real) <= 1e-9
    assert resid.max() <= 1e-7
