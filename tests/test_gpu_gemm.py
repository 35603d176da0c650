"""MFMA GEMM family through the C ABI (sisi_gemm_host) against numpy FP64."""

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu

# error bound relative to max_{mn} sum_k |A_mk||B_kn|
TOL = {0: 1e-14, 1: 4e-7, 2: 4e-5, 3: 2e-2}


def _run(prec, ta, tb, M, N, K, alpha=1.0, beta=0.0, seed=0):
    from sisi4s_amd import engine
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    announce("gemm prec=%d ta=%d tb=%d M=%d N=%d K=%d" % (prec, ta, tb, M, N, K))
    C = engine.gemm(A, B, C0, alpha=alpha, beta=beta, transA=ta, transB=tb, precision=prec)
    Aop = A.T if ta else A
    Bop = B.T if tb else B
    ref = alpha * Aop @ Bop + beta * C0
    bound = np.abs(alpha) * (np.abs(Aop) @ np.abs(Bop)).max() + abs(beta) * np.abs(C0).max()
    err = np.abs(C - ref).max()
    assert err <= TOL[prec] * max(bound, 1.0) * max(1.0, np.sqrt(K) / 8), (err, bound)


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 5), (130, 129, 70), (256, 384, 96)])
def test_gemm_shapes(prec, ta, tb, M, N, K):
    _run(prec, ta, tb, M, N, K, seed=M + N + K)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_gemm_alpha_beta_and_empty_k(prec):
    _run(prec, 0, 0, 40, 50, 60, alpha=-0.5, beta=2.0)
    _run(prec, 1, 0, 40, 50, 0, alpha=1.0, beta=-3.0)


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_gemm_split_k(prec):
    # few output tiles and a long k: the launcher splits k and reduces in FP64
    _run(prec, 1, 0, 64, 48, 40000, seed=3)


@pytest.mark.parametrize("prec", [2, 3])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(17, 33, 5), (256, 384, 96), (300, 200, 3000)])
def test_gemm_library_backend(prec, ta, tb, M, N, K):
    """The hipBLASLt backend of the split-BF16x3 / bf16 products (operand
    planes with K tripled, FP64 epilogue) meets the same error bounds."""
    from sisi4s_amd import engine
    engine.set_bf16_backend(True)
    try:
        _run(prec, ta, tb, M, N, K, alpha=0.75, beta=0.5, seed=M + K)
    finally:
        engine.set_bf16_backend(False)


def _variant_ids(passes):
    from sisi4s_amd import engine
    return list(range(len(engine.gemm_variants(passes))))


@pytest.mark.parametrize("passes,prec", [(3, 2), (1, 3)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(17, 33, 5), (256, 384, 96), (300, 200, 3000),
                                   (520, 270, 700)])
def test_gemm_every_kernel_variant(passes, prec, ta, tb, M, N, K):
    """Every kernel variant of the split-BF16x3 and bf16 products (the
    128-row register-staged kernels and the 256-row direct-to-LDS kernels,
    with their split-K paths) meets the precision's error bound."""
    from sisi4s_amd import engine
    for v in _variant_ids(passes):
        engine.set_gemm_variant(passes, v)
        try:
            _run(prec, ta, tb, M, N, K, alpha=0.75, beta=0.5, seed=M + N + K + v)
        finally:
            engine.set_gemm_variant(passes, -1)


def test_gemm_bench_and_tuning_cache():
    """The timing entry point returns positive times, and a tuning pass on a
    fresh cache records one winner for the class it met."""
    from sisi4s_amd import engine
    announce("gemm bench + tuning")
    k, t = engine.gemm_bench("bf16x3", 1024, 1024, 512, reps=2)
    assert 0.0 < k <= t
    engine.tune_enable(True)
    try:
        engine.gemm_bench("bf16", 2048, 1536, 256, reps=1)
    finally:
        engine.tune_enable(False)
    dump = engine.tune_dump()
    assert any(line.startswith("1 11 11 8 ") for line in dump.splitlines()), dump
