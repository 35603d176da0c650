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
    # SISI_SANITIZE=1: the harness (and the host headers it includes) built
    # with AddressSanitizer and UndefinedBehaviorSanitizer
    san = (["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=undefined"] if os.environ.get("SISI_SANITIZE") == "1" else ["-O2"])
    subprocess.run([gxx] + san + ["-std=c++17", "-Wall", "-Werror", SRC, "-o", exe], check=True)
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


def _complex_lowest_pair(n, seed):
    """A non-symmetric matrix whose two lowest eigenvalues (by real part) are
    the complex pair -2 +- 0.05i (a 2 x 2 rotation block), the others real and
    well separated above, under a similarity close to the identity so that
    the diagonal is a useful preconditioner."""
    rng = np.random.default_rng(seed)
    B = np.diag(np.linspace(-1.0, 6.0, n))
    B[:2, :2] = [[-2.0, 0.05], [-0.05, -2.0]]
    S = np.eye(n) + 0.01 * rng.normal(size=(n, n))
    A = S @ B @ np.linalg.inv(S)
    ref = np.sort_complex(np.linalg.eigvals(A))
    assert abs(ref[0] - (-2.0 - 0.05j)) <= 1e-10 and abs(ref[1] - (-2.0 + 0.05j)) <= 1e-10
    return A, np.diag(A).copy()


@pytest.mark.parametrize("refresh,max_basis", [(0, 1000), (1, 6)])
def test_davidson_complex_lowest_pair(harness, refresh, max_basis):
    """A complex-conjugate lowest pair: the two roots must come from the
    pair's two-dimensional invariant subspace (Re and Im of the complex Ritz
    vector), not from one real shift used twice, with and without restarts
    on a small basis."""
    A, d = _complex_lowest_pair(60, 21)
    vals, resid, info = davidson(harness, A, d, 2, tol_vectors=1e-22, tol_energy=0.0,
                                 max_basis=max_basis, max_iterations=400, refresh=refresh)
    assert info["converged"], info
    assert abs(vals[0].real + 2.0) <= 1e-9 and abs(vals[1].real + 2.0) <= 1e-9
    assert abs(abs(vals[0].imag) - 0.05) <= 1e-9 and abs(vals[0].imag + vals[1].imag) <= 1e-12
    assert resid.max() <= 1e-9


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


def _lines(exe, text):
    return [np.array([float(x) for x in l.split()]) for l in _run(exe, text)]


def _fmt(x):
    return " ".join("%.17g" % y for y in np.asarray(x).ravel(order="F"))


@pytest.mark.parametrize("ea", [False, True])
def test_eom_space_packing_reproduces_the_oracle_spectrum(harness, ea):
    """The EOM steps' packed space (sisi4s_amd/host/eom_space.hpp): unpacking
    the unit vectors, applying the oracle's sigma to the dense r1 / r2 arrays
    in the layout the C ABI takes, and packing the result gives a matrix with
    the oracle's IP / EA spectrum; unpack then pack is the identity."""
    from oracle import eom
    from test_eom_program import _system
    ints, ei, ea_, T1, T2, _ = _system(3, 8, 31)
    o, v = 3, 5
    n1, r2shape = (v, (o, v, v)) if ea else (o, (o, o, v))
    n = n1 + (v * (v - 1) // 2 * o if ea else o * (o - 1) // 2 * v)
    head = "%d %d %d" % (int(ea), o, v)
    dense = _lines(harness, "unpack %s %d\n%s\n" % (head, n, _fmt(np.eye(n))))
    assert len(dense) == n
    H = (eom.hbar_ea if ea else eom.hbar_ip)(ints, ei, ea_, T1, T2)
    sigma = eom.ea_sigma if ea else eom.ip_sigma
    cols = []
    for x in dense:
        r1, r2 = x[:n1], x[n1:].reshape(r2shape, order="F")
        assert np.abs(r2 + (r2.transpose(0, 2, 1) if ea else r2.transpose(1, 0, 2))).max() == 0.0
        s1, s2 = sigma(H, r1, r2)
        cols.append(_fmt(s1) + " " + _fmt(s2))
    M = np.array(_lines(harness, "pack %s %d\n%s\n" % (head, n, "\n".join(cols)))).T
    ref = (eom.ea_eigenvalues if ea else eom.ip_eigenvalues)(ints, ei, ea_, T1, T2)
    assert np.abs(np.sort(np.linalg.eigvals(M).real) - ref).max() <= 1e-11
    back = np.array(_lines(harness, "pack %s %d\n%s\n" % (head, n, "\n".join(_fmt(x) for x in dense))))
    assert np.array_equal(back, np.eye(n))
    d = _lines(harness, "diag %s\n%s %s\n" % (head, _fmt(ei), _fmt(ea_)))[0]
    # the diagonal of the zeroth-order (Fock) part of Hbar: sigma with T = 0 and no integrals
    zero = {k: np.zeros_like(x) for k, x in ints.items()}
    H0 = (eom.hbar_ea if ea else eom.hbar_ip)(zero, ei, ea_, np.zeros_like(T1), np.zeros_like(T2))
    D = []
    for x in dense:
        s1, s2 = sigma(H0, x[:n1], x[n1:].reshape(r2shape, order="F"))
        D.append(np.concatenate([s1, s2.ravel(order="F")]) @ x / (x @ x))
    # each unit vector x is an eigenvector of H0; its Rayleigh quotient in the
    # dense metric equals the packed diagonal element
    assert np.abs(np.array(D) - d).max() <= 1e-14


def test_class_reorderings_of_the_eom_steps(harness):
    """HHPP -> PPHH, HPPP -> PPPH, HPHP -> PHPH as the steps upload them,
    against slices of one antisymmetrized tensor."""
    rng = np.random.default_rng(5)
    n, o = 7, 3
    g = rng.normal(size=(6, n, n))
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    a = eri - np.swapaxes(eri, 2, 3)
    h, p = slice(0, o), slice(o, n)
    v = n - o
    out = _lines(harness, "classes %d %d\n%s\n%s\n%s\n" % (o, v, _fmt(a[h, h, p, p]), _fmt(a[h, p, p, p]),
                                                          _fmt(a[h, p, h, p])))
    assert np.abs(out[0] - a[p, p, h, h].ravel(order="F")).max() == 0.0
    assert np.abs(out[1] - a[p, p, p, h].ravel(order="F")).max() <= 1e-15
    assert np.abs(out[2] - a[p, h, p, h].ravel(order="F")).max() == 0.0
