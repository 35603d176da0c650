"""ctypes binding of the C ABI in include/sisi_engine.h.

Arrays cross the boundary as float64 buffers in the reference's CTF layout
(first index fastest).  Python callers pass and receive logical tensors
(T2[a,b,i,j] etc.); the conversion to the buffer layout is done here.

The binding refuses to run without the compiled library: there is no CPU
fallback for any entry point.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsisi_engine.so")

SISI_OK = 0
SISI_ERR_INVALID = 1
SISI_ERR_HIP = 2
SISI_ERR_STATE = 3
SISI_ERR_NUMERIC = 4
SISI_ERR_ALLOC = 5

FP64, FP32, BF16X3, BF16 = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "fp32": FP32, "bf16x3": BF16X3, "bf16": BF16}
MIXER_LINEAR, MIXER_DIIS = 0, 1
PPHH, PHPH, HHHH, HHHP, HHPP, PHHP, PHHH, PPPP, PPPH = 0, 1, 2, 3, 4, 5, 6, 7, 8
FORMULATION_VERTEX, FORMULATION_INTEGRALS, FORMULATION_ONLY_PPL, FORMULATION_DRCCD = 0, 1, 2, 3
FORMULATION_UCCSD = 4

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int_p = ctypes.POINTER(ctypes.c_int)


class Dims(ctypes.Structure):
    _fields_ = [("No", ctypes.c_int), ("Nv", ctypes.c_int), ("NG", ctypes.c_int)]


class CcsdOpts(ctypes.Structure):
    _fields_ = [("mixer", ctypes.c_int), ("max_residua", ctypes.c_int),
                ("mixing_ratio", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("amplitudes_convergence", ctypes.c_double),
                ("energy_convergence", ctypes.c_double), ("level_shift", ctypes.c_double),
                ("slice_size", ctypes.c_int), ("precision", ctypes.c_int),
                ("switch_threshold", ctypes.c_double), ("refine_interval", ctypes.c_int),
                ("refine_precision", ctypes.c_int), ("ppl", ctypes.c_int),
                ("distinguishable", ctypes.c_int), ("coarse_threshold", ctypes.c_double),
                ("coarse_precision", ctypes.c_int), ("graphs", ctypes.c_int),
                ("bulk_refine_interval", ctypes.c_int)]


class CcsdReport(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("direct", ctypes.c_double),
                ("exchange", ctypes.c_double), ("delta", ctypes.c_double),
                ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("fp64_iterations", ctypes.c_int), ("precision", ctypes.c_int),
                ("status", ctypes.c_int), ("energy_imag", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LockstepStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_double), ("launches", ctypes.c_double),
                ("records", ctypes.c_double), ("plans", ctypes.c_double),
                ("setup_seconds", ctypes.c_double), ("residual_seconds", ctypes.c_double),
                ("update_seconds", ctypes.c_double), ("total_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("gemm_seconds", ctypes.c_double * 3), ("gemm_flops", ctypes.c_double * 3),
                ("gemm_launches", ctypes.c_int64 * 3), ("other_seconds", ctypes.c_double),
                ("other_bytes", ctypes.c_double), ("other_launches", ctypes.c_int64)]

    def as_dict(self):
        return dict(gemm_seconds=list(self.gemm_seconds), gemm_flops=list(self.gemm_flops),
                    gemm_launches=list(self.gemm_launches), other_seconds=self.other_seconds,
                    other_bytes=self.other_bytes, other_launches=self.other_launches)


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sisi engine error %d: %s" % (code, msg))
        self.code = code


# (name, restype, argtypes) of every exported function
SIGNATURES = [
    ("sisi_last_error", ctypes.c_char_p, []),
    ("sisi_ccsd_opts_default", None, [ctypes.POINTER(CcsdOpts)]),
    ("sisi_engine_create", ctypes.c_int, [ctypes.POINTER(Dims), ctypes.POINTER(ctypes.c_void_p)]),
    ("sisi_engine_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("sisi_engine_set_vertex", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("sisi_engine_set_eigenenergies", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    ("sisi_engine_set_vertex_complex", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("sisi_integrals_from_vertex", ctypes.c_int, [ctypes.c_void_p]),
    ("sisi_get_integrals", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_double_p]),
    ("sisi_engine_set_precision", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("sisi_engine_set_residual_variant", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("sisi_engine_set_integrals", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_double_p]),
    ("sisi_set_bf16_backend", ctypes.c_int, [ctypes.c_int]),
    ("sisi_engine_set_formulation", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("sisi_engine_set_drccd", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("sisi_eom_set_amplitudes", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    ("sisi_eom_ip_sigma", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p, _c_double_p,
                                         _c_double_p, _c_double_p, _c_double_p]),
    ("sisi_eom_ea_sigma", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p, _c_double_p,
                                         _c_double_p, _c_double_p, _c_double_p]),
    ("sisi_engine_set_coulomb_factors", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_double_p, _c_double_p, _c_double_p]),
    ("sisi_ccsd_residual", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_double_p, _c_double_p,
      _c_double_p, _c_double_p]),
    ("sisi_ccsd_residual_anchor", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, _c_double_p, _c_double_p, _c_double_p, _c_double_p]),
    ("sisi_ccsd_residual_refine", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, _c_double_p, _c_double_p, _c_double_p, _c_double_p]),
    ("sisi_ccsd_energy", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p, _c_double_p]),
    ("sisi_mp2_energy", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("sisi_ccsd_begin", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(CcsdOpts), _c_double_p, _c_double_p]),
    ("sisi_ccsd_iterate", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(CcsdReport)]),
    ("sisi_ccsd_solve", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(CcsdOpts), _c_double_p, _c_double_p,
      ctypes.POINTER(CcsdReport)]),
    ("sisi_ccsd_get_amplitudes", ctypes.c_int, [ctypes.c_void_p, _c_double_p, _c_double_p]),
    ("sisi_ccsd_get_energies", ctypes.c_int, [ctypes.c_void_p, _c_double_p, ctypes.c_int, _c_int_p]),
    ("sisi_ccsd_pair_energies", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("sisi_triples_energy", ctypes.c_int,
     [ctypes.c_void_p, _c_double_p, _c_double_p, ctypes.c_int, _c_double_p]),
    ("sisi_triples_energy_blocks", ctypes.c_int,
     [ctypes.c_void_p, _c_double_p, _c_double_p, ctypes.c_int, _c_int_p, ctypes.c_int,
      _c_double_p]),
    ("sisi_ccsd_solve_batched", ctypes.c_int,
     [ctypes.c_int, ctypes.POINTER(Dims), ctypes.POINTER(_c_double_p), ctypes.POINTER(_c_double_p),
      ctypes.POINTER(_c_double_p), ctypes.POINTER(CcsdOpts), ctypes.c_int, _c_double_p,
      ctypes.POINTER(CcsdReport)]),
    ("sisi_engine_set_ladder_planes", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("sisi_ccsd_solve_lockstep", ctypes.c_int,
     [ctypes.c_int, ctypes.POINTER(Dims), ctypes.POINTER(_c_double_p), ctypes.POINTER(_c_double_p),
      ctypes.POINTER(_c_double_p), ctypes.POINTER(CcsdOpts), _c_double_p,
      ctypes.POINTER(CcsdReport), ctypes.POINTER(LockstepStats)]),
    ("sisi_engine_enable_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("sisi_engine_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("sisi_engine_reset_stats", ctypes.c_int, [ctypes.c_void_p]),
    ("sisi_generate_system", ctypes.c_int,
     [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_double_p, _c_double_p,
      _c_double_p]),
    ("sisi_generate_ueg", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, _c_int_p,
      _c_double_p, _c_double_p, _c_double_p]),
    ("sisi_gemm_host", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
      ctypes.c_double, _c_double_p, ctypes.c_int, _c_double_p, ctypes.c_int, ctypes.c_double,
      _c_double_p, ctypes.c_int]),
    ("sisi_engine_set_energy_variant", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("sisi_engine_set_hp_fock", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("sisi_engine_get_noncanonical_energy", ctypes.c_int, [ctypes.c_void_p, _c_double_p]),
    ("sisi_gemm_set_variant", ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    ("sisi_gemm_variant_count", ctypes.c_int, [ctypes.c_int]),
    ("sisi_gemm_variant_name", ctypes.c_char_p, [ctypes.c_int, ctypes.c_int]),
    ("sisi_tune_enable", ctypes.c_int, [ctypes.c_int]),
    ("sisi_tune_dump", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _c_int_p]),
    ("sisi_gemm_device_noalloc", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
      ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    ("sisi_gemm_bench", ctypes.c_int,
     [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_double_p,
      _c_double_p]),
]

_lib = None


def load(path=None):
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            "HIP engine library %s is missing: build it with `python -m sisi4s_amd.build`" % p)
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc):
    if rc != SISI_OK:
        msg = load().sisi_last_error()
        raise EngineError(rc, msg.decode() if msg else "")


def _buf(x):
    """Logical float64 tensor -> contiguous first-index-fastest buffer."""
    b = np.asarray(x, dtype=np.float64).ravel(order="F")
    return np.ascontiguousarray(b)


def _ptr(b):
    return b.ctypes.data_as(_c_double_p) if b is not None else None


def _sized(x, n, name, cplx=False):
    """_buf(x) after checking that x holds exactly n logical elements, so
    the library never reads past a caller's array."""
    if np.size(x) != n:
        raise ValueError("%s: expected %d elements, got %d (shape %s)"
                         % (name, n, np.size(x), np.shape(x)))
    if cplx:
        b = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).ravel(order="F"))
        return b.view(np.float64)
    return _buf(x)


def _unbuf(b, shape):
    return b.reshape(shape, order="F")


def set_bf16_backend(library):
    """Process-wide backend of the split-BF16x3 / bf16 products: False = the
    engine's MFMA kernel, True = hipBLASLt on the engine's operand planes."""
    _check(load().sisi_set_bf16_backend(int(bool(library))))


def set_gemm_variant(passes, variant):
    """Force kernel variant `variant` of the split-BF16x3 (passes=3) or bf16
    (passes=1) products for the whole process; -1 restores the automatic
    choice (tuning cache, else the size rule)."""
    _check(load().sisi_gemm_set_variant(int(passes), int(variant)))


def gemm_variants(passes):
    """Names of the kernel variants of a pass count, by variant id."""
    lib = load()
    return [lib.sisi_gemm_variant_name(int(passes), v).decode()
            for v in range(lib.sisi_gemm_variant_count(int(passes)))]


def tune_enable(on=True):
    """Time unknown product classes on first use and record the fastest
    variant (process-wide)."""
    _check(load().sisi_tune_enable(int(bool(on))))


def tune_dump():
    """The tuning records of this process as text lines."""
    lib = load()
    n = ctypes.c_int(0)
    _check(lib.sisi_tune_dump(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(max(n.value, 1))
    _check(lib.sisi_tune_dump(buf, len(buf), ctypes.byref(n)))
    return buf.value.decode()


def gemm_bench(precision, M, N, K, reps=5):
    """(kernel microseconds, whole-call microseconds) per GEMM call on
    device-resident operands of the given shape."""
    if isinstance(precision, str):
        precision = PRECISIONS[precision]
    k, t = ctypes.c_double(), ctypes.c_double()
    _check(load().sisi_gemm_bench(int(precision), int(M), int(N), int(K), int(reps),
                                  ctypes.byref(k), ctypes.byref(t)))
    return k.value, t.value


def default_opts(**kw):
    o = CcsdOpts()
    load().sisi_ccsd_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if k == "mixer" and isinstance(v, str):
            v = {"LinearMixer": MIXER_LINEAR, "DiisMixer": MIXER_DIIS}[v]
        if k in ("precision", "refine_precision", "coarse_precision") and isinstance(v, str):
            v = PRECISIONS[v]
        setattr(o, k, v)
    return o


def generate_system(seed, No, Nv, NG):
    """Seeded synthetic (Gamma[G,p,q], eps_i, eps_a) from the engine's
    host generator (no device needed)."""
    Np = No + Nv
    g = np.empty(NG * Np * Np)
    ei = np.empty(No)
    ea = np.empty(Nv)
    _check(load().sisi_generate_system(int(seed), No, Nv, NG, _ptr(g), _ptr(ei), _ptr(ea)))
    return _unbuf(g, (NG, Np, Np)), ei, ea


def generate_ueg(No, Nv, rs, hartree_fock=True, madelung=-1.0):
    """Uniform electron gas (Gamma[G,p,q] in real orbitals, eps_i, eps_a)
    from the engine's host generator (no device needed)."""
    lib = load()
    ng = ctypes.c_int(0)
    _check(lib.sisi_generate_ueg(No, Nv, float(rs), int(bool(hartree_fock)), float(madelung),
                                 ctypes.byref(ng), None, None, None))
    Np = No + Nv
    g = np.empty(ng.value * Np * Np)
    ei = np.empty(No)
    ea = np.empty(Nv)
    _check(lib.sisi_generate_ueg(No, Nv, float(rs), int(bool(hartree_fock)), float(madelung),
                                 ctypes.byref(ng), _ptr(g), _ptr(ei), _ptr(ea)))
    return _unbuf(g, (ng.value, Np, Np)), ei, ea


def gemm(A, B, C=None, alpha=1.0, beta=0.0, transA=False, transB=False, precision=FP64):
    """C = alpha op(A) op(B) + beta C through one MFMA GEMM launch."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    M = A.shape[1] if transA else A.shape[0]
    K = A.shape[0] if transA else A.shape[1]
    N = B.shape[0] if transB else B.shape[1]
    Cb = np.zeros((M, N), order="F") if C is None else np.array(C, dtype=np.float64, order="F")
    _check(load().sisi_gemm_host(int(precision), int(transA), int(transB), M, N, K, alpha,
                                 A.ctypes.data_as(_c_double_p), max(A.shape[0], 1),
                                 B.ctypes.data_as(_c_double_p), max(B.shape[0], 1), beta,
                                 Cb.ctypes.data_as(_c_double_p), max(M, 1)))
    return Cb


def _batch_inputs(systems):
    n = len(systems)
    dims = (Dims * max(n, 1))()
    bufs = []
    gp = (_c_double_p * max(n, 1))()
    ip = (_c_double_p * max(n, 1))()
    ap = (_c_double_p * max(n, 1))()
    for s, (g, ei, ea) in enumerate(systems):
        No, Nv = len(ei), len(ea)
        NG = int(np.shape(g)[0])
        dims[s] = Dims(No, Nv, NG)
        b = (_sized(g, NG * (No + Nv) ** 2, "system %d gamma[NG,Np,Np]" % s), _buf(ei), _buf(ea))
        bufs.append(b)
        gp[s], ip[s], ap[s] = _ptr(b[0]), _ptr(b[1]), _ptr(b[2])
    return dims, bufs, gp, ip, ap


def ccsd_solve_lockstep(systems, opts=None, strict=True):
    """The batch of ccsd_solve_batched iterated in lockstep: every launch of
    an iteration is one grouped launch over all systems.  Returns (energies,
    reports, stats) with stats the launch counters of the run."""
    n = len(systems)
    dims, bufs, gp, ip, ap = _batch_inputs(systems)
    o = opts if opts is not None else default_opts()
    energies = np.zeros(max(n, 1))
    reps = (CcsdReport * max(n, 1))()
    st = LockstepStats()
    rc = load().sisi_ccsd_solve_lockstep(n, dims, gp, ip, ap, ctypes.byref(o), _ptr(energies),
                                         reps, ctypes.byref(st))
    del bufs
    if strict:
        _check(rc)
    return energies[:n].copy(), [reps[s].as_dict() for s in range(n)], st.as_dict()


def ccsd_solve_batched(systems, opts=None, workers=8, strict=True):
    """CCSD on many independent systems [(gamma, eps_i, eps_a), ...] in one
    call; returns (energies, reports).  With strict=False a failed system
    gets NaN and its status in reports[s]["status"] instead of an error."""
    n = len(systems)
    dims = (Dims * max(n, 1))()
    bufs = []
    gp = (_c_double_p * max(n, 1))()
    ip = (_c_double_p * max(n, 1))()
    ap = (_c_double_p * max(n, 1))()
    for s, (g, ei, ea) in enumerate(systems):
        No, Nv = len(ei), len(ea)
        NG = int(np.shape(g)[0])
        dims[s] = Dims(No, Nv, NG)
        b = (_sized(g, NG * (No + Nv) ** 2, "system %d gamma[NG,Np,Np]" % s), _buf(ei), _buf(ea))
        bufs.append(b)
        gp[s], ip[s], ap[s] = _ptr(b[0]), _ptr(b[1]), _ptr(b[2])
    o = opts if opts is not None else default_opts()
    energies = np.zeros(max(n, 1))
    reps = (CcsdReport * max(n, 1))()
    rc = load().sisi_ccsd_solve_batched(n, dims, gp, ip, ap, ctypes.byref(o), int(workers),
                                        _ptr(energies), reps)
    if strict:
        _check(rc)
    return energies[:n].copy(), [reps[s].as_dict() for s in range(n)]


class Engine:
    """One engine handle = device copies of one system + one HIP stream."""

    def __init__(self, No, Nv, NG):
        self.No, self.Nv, self.NG = No, Nv, NG
        self.Np = No + Nv
        self._lib = load()
        h = ctypes.c_void_p()
        _check(self._lib.sisi_engine_create(ctypes.byref(Dims(No, Nv, NG)), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sisi_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- tensors at the boundary (complex128 interleaved in complex mode) --
    complex = False

    def _in(self, x, n=None, name="tensor"):
        if n is not None:
            return _sized(x, n, name, self.complex)
        if not self.complex:
            return _buf(x)
        b = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).ravel(order="F"))
        return b.view(np.float64)

    def _t1(self, T1):
        return self._in(T1, self.Nv * self.No, "T1[Nv,No]")

    def _t2(self, T2):
        return self._in(T2, self.Nv * self.Nv * self.No * self.No, "T2[Nv,Nv,No,No]")

    def _out(self, shape):
        n = int(np.prod(shape))
        return np.empty(2 * n if self.complex else n)

    def _shape(self, b, shape):
        if self.complex:
            return b.view(np.complex128).reshape(shape, order="F")
        return _unbuf(b, shape)

    # --- inputs -----------------------------------------------------------
    def set_vertex(self, gamma):
        g = _sized(gamma, self.NG * self.Np * self.Np, "gamma[NG,Np,Np]")
        _check(self._lib.sisi_engine_set_vertex(self._h, _ptr(g)))
        self.complex = False

    def set_vertex_complex(self, gamma):
        """Complex vertex: the engine switches to the reference's complex
        path (CcsdEnergyFromCoulombIntegrals.cxx:496-891)."""
        g = _sized(gamma, self.NG * self.Np * self.Np, "gamma[NG,Np,Np]", cplx=True)
        _check(self._lib.sisi_engine_set_vertex_complex(self._h, _ptr(g)))
        self.complex = True

    def set_eigenenergies(self, eps_i, eps_a):
        ei = _sized(eps_i, self.No, "eps_i[No]")
        ea = _sized(eps_a, self.Nv, "eps_a[Nv]")
        _check(self._lib.sisi_engine_set_eigenenergies(self._h, _ptr(ei), _ptr(ea)))

    def set_system(self, gamma, eps_i, eps_a, integrals=True):
        if np.iscomplexobj(gamma):
            self.set_vertex_complex(gamma)
        else:
            self.set_vertex(gamma)
        self.set_eigenenergies(eps_i, eps_a)
        if integrals:
            self.integrals_from_vertex()

    def integrals_from_vertex(self):
        _check(self._lib.sisi_integrals_from_vertex(self._h))

    def get_integrals(self, which):
        o, v = self.No, self.Nv
        shapes = {PPHH: (v, v, o, o), PHPH: (v, o, v, o), HHHH: (o, o, o, o),
                  HHHP: (o, o, o, v), HHPP: (o, o, v, v), PHHP: (v, o, o, v),
                  PHHH: (v, o, o, o), PPPP: (v, v, v, v), PPPH: (v, v, v, o)}
        shape = shapes[which]
        b = self._out(shape)
        _check(self._lib.sisi_get_integrals(self._h, which, _ptr(b)))
        return self._shape(b, shape)

    def set_precision(self, precision):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        _check(self._lib.sisi_engine_set_precision(self._h, int(precision)))

    def set_coulomb_factors(self, Lambda=None, Pi=None, slice_size=-1, PiOut=None):
        """CoulombFactors Lambda[G,R], FactorOrbitals Pi[p,R] and (complex
        amplitudes) OutgoingFactorOrbitals PiOut[p,R]: the ladder becomes the
        factor-based one (CcsdEnergyFromCoulombIntegrals.cxx:340-372,
        786-813); no arguments remove them."""
        if Lambda is None:
            _check(self._lib.sisi_engine_set_coulomb_factors(self._h, 0, 0, None, None, None))
            return

        def z(a):
            return np.ascontiguousarray(np.asarray(a, dtype=np.complex128).ravel(order="F"))
        L, P = z(Lambda), z(Pi)
        O = z(PiOut) if PiOut is not None else None
        NR = int(np.shape(Lambda)[1])
        if L.size != self.NG * NR or P.size != self.Np * NR or (O is not None and O.size != P.size):
            raise ValueError("set_coulomb_factors: Lambda must be [NG,NR], Pi and PiOut [Np,NR]")
        _check(self._lib.sisi_engine_set_coulomb_factors(
            self._h, NR, int(slice_size), _ptr(L.view(np.float64)), _ptr(P.view(np.float64)),
            _ptr(O.view(np.float64)) if O is not None else None))

    def set_integrals(self, which, V):
        """Upload one real integral class (no vertex needed for the
        full-integral formulation)."""
        o, v = self.No, self.Nv
        sizes = {PPHH: v * v * o * o, PHPH: v * o * v * o, HHHH: o ** 4, HHHP: o ** 3 * v,
                 HHPP: o * o * v * v, PHHP: v * o * o * v, PHHH: v * o ** 3, PPPP: v ** 4,
                 PPPH: v ** 3 * o}
        if int(which) not in sizes:
            raise ValueError("set_integrals: unknown integral class %r" % (which,))
        b = _sized(V, sizes[int(which)], "integrals class %d" % int(which))
        _check(self._lib.sisi_engine_set_integrals(self._h, int(which), _ptr(b)))

    def set_formulation(self, formulation):
        _check(self._lib.sisi_engine_set_formulation(self._h, int(formulation)))

    def set_drccd(self, unrestricted=False, linearized=False, adjacent_pairs_exchange=False):
        """Direct-ring CCD residual (DrccdEnergyFromCoulombIntegrals.cxx:31-80)
        for the next residual calls and solves, with its arguments."""
        _check(self._lib.sisi_engine_set_drccd(self._h, int(bool(unrestricted)), int(bool(linearized)),
                                               int(bool(adjacent_pairs_exchange))))
        self.set_formulation(FORMULATION_DRCCD)

    def eom_ip_sigma(self, T1, T2, r1, r2):
        """IP-EOM-CCSD right-apply (SimilarityTransformedHamiltonian.cxx:64-68)
        on a spin-orbital engine: returns (s1[i], s2[i,j,a])."""
        o, v = self.No, self.Nv
        b1, b2 = np.empty(o), np.empty(o * o * v)
        t1 = _sized(T1, v * o, "T1[Nv,No]")
        t2 = _sized(T2, v * v * o * o, "T2[Nv,Nv,No,No]")
        x1, x2 = _sized(r1, o, "r1[No]"), _sized(r2, o * o * v, "r2[No,No,Nv]")
        _check(self._lib.sisi_eom_ip_sigma(self._h, _ptr(t1), _ptr(t2), _ptr(x1), _ptr(x2),
                                           _ptr(b1), _ptr(b2)))
        return _unbuf(b1, (o,)), _unbuf(b2, (o, o, v))

    def eom_ea_sigma(self, T1, T2, r1, r2):
        """EA-EOM-CCSD right-apply (SimilarityTransformedHamiltonian.cxx:289-294)
        on a spin-orbital engine with PPPP uploaded: returns (s1[a], s2[j,a,b])."""
        o, v = self.No, self.Nv
        b1, b2 = np.empty(v), np.empty(o * v * v)
        t1 = _sized(T1, v * o, "T1[Nv,No]")
        t2 = _sized(T2, v * v * o * o, "T2[Nv,Nv,No,No]")
        x1, x2 = _sized(r1, v, "r1[Nv]"), _sized(r2, o * v * v, "r2[No,Nv,Nv]")
        _check(self._lib.sisi_eom_ea_sigma(self._h, _ptr(t1), _ptr(t2), _ptr(x1), _ptr(x2),
                                           _ptr(b1), _ptr(b2)))
        return _unbuf(b1, (v,)), _unbuf(b2, (o, v, v))

    def set_ladder_planes(self, on=True):
        """Reduced-precision ladder with plane output (default on): the V~
        slice-pair product writes the bf16 planes the X contraction reads."""
        _check(self._lib.sisi_engine_set_ladder_planes(self._h, int(bool(on))))

    def set_energy_variant(self, unrestricted=False, antisymmetrize=False):
        """`unrestricted` (spins = 1) and `antisymmetrize` energies of getEnergy
        (ClusterSinglesDoublesAlgorithm.cxx:131-226)."""
        _check(self._lib.sisi_engine_set_energy_variant(self._h, int(bool(unrestricted)),
                                                        int(bool(antisymmetrize))))

    def set_hp_fock(self, fia=None):
        """HPFockMatrix f[i,a] (No x Nv): energies add spins * sum T_ai f_ia
        (ClusterSinglesDoublesAlgorithm.cxx:199-211); None clears it."""
        if fia is None:
            _check(self._lib.sisi_engine_set_hp_fock(self._h, None))
            return
        b = self._in(fia, self.No * self.Nv, "HPFockMatrix[No,Nv]")
        _check(self._lib.sisi_engine_set_hp_fock(self._h, _ptr(b)))

    def noncanonical_energy(self):
        out = ctypes.c_double()
        _check(self._lib.sisi_engine_get_noncanonical_energy(self._h, ctypes.byref(out)))
        return out.value

    def set_residual_variant(self, ppl=True, distinguishable=False):
        """ppl / distinguishable of the next residual calls
        (CcsdEnergyFromCoulombIntegrals.cxx:123-124, 324-338)."""
        _check(self._lib.sisi_engine_set_residual_variant(self._h, int(bool(ppl)),
                                                          int(bool(distinguishable))))

    # --- residual / energies ---------------------------------------------
    def residual(self, iteration, T1, T2, slice_size=-1, initial_doubles=False):
        v, o = self.Nv, self.No
        t1, t2 = self._t1(T1), self._t2(T2)
        r1, r2 = self._out((v, o)), self._out((v, v, o, o))
        _check(self._lib.sisi_ccsd_residual(self._h, iteration, int(initial_doubles),
                                            int(slice_size), _ptr(t1), _ptr(t2), _ptr(r1),
                                            _ptr(r2)))
        return self._shape(r1, (v, o)), self._shape(r2, (v, v, o, o))

    def residual_anchor(self, T1, T2, slice_size=-1):
        return self._residual2(self._lib.sisi_ccsd_residual_anchor, T1, T2, slice_size)

    def residual_refine(self, T1, T2, slice_size=-1):
        return self._residual2(self._lib.sisi_ccsd_residual_refine, T1, T2, slice_size)

    def _residual2(self, fn, T1, T2, slice_size):
        v, o = self.Nv, self.No
        t1 = _sized(T1, v * o, "T1[Nv,No]")
        t2 = _sized(T2, v * v * o * o, "T2[Nv,Nv,No,No]")
        r1 = np.empty(v * o)
        r2 = np.empty(v * v * o * o)
        _check(fn(self._h, int(slice_size), _ptr(t1), _ptr(t2), _ptr(r1), _ptr(r2)))
        return _unbuf(r1, (v, o)), _unbuf(r2, (v, v, o, o))

    def energy(self, T1, T2):
        out = np.empty(3)
        _check(self._lib.sisi_ccsd_energy(self._h, _ptr(self._t1(T1)), _ptr(self._t2(T2)),
                                          _ptr(out)))
        return tuple(out)

    def mp2_energy(self):
        e = ctypes.c_double()
        _check(self._lib.sisi_mp2_energy(self._h, ctypes.byref(e)))
        return e.value

    # --- driver ------------------------------------------------------------
    def ccsd_begin(self, opts=None, T1=None, T2=None):
        o = opts if opts is not None else default_opts()
        t1 = self._t1(T1) if T1 is not None else None
        t2 = self._t2(T2) if T2 is not None else None
        _check(self._lib.sisi_ccsd_begin(self._h, ctypes.byref(o), _ptr(t1), _ptr(t2)))

    def ccsd_iterate(self, n):
        r = CcsdReport()
        _check(self._lib.sisi_ccsd_iterate(self._h, int(n), ctypes.byref(r)))
        return r.as_dict()

    def ccsd_solve(self, opts=None, T1=None, T2=None):
        o = opts if opts is not None else default_opts()
        t1 = self._t1(T1) if T1 is not None else None
        t2 = self._t2(T2) if T2 is not None else None
        r = CcsdReport()
        _check(self._lib.sisi_ccsd_solve(self._h, ctypes.byref(o), _ptr(t1), _ptr(t2),
                                         ctypes.byref(r)))
        return r.as_dict()

    def amplitudes(self):
        v, o = self.Nv, self.No
        t1, t2 = self._out((v, o)), self._out((v, v, o, o))
        _check(self._lib.sisi_ccsd_get_amplitudes(self._h, _ptr(t1), _ptr(t2)))
        return self._shape(t1, (v, o)), self._shape(t2, (v, v, o, o))

    def pair_energies(self):
        """PairEnergy [No, No] of the current amplitudes."""
        p = self._out((self.No, self.No))
        _check(self._lib.sisi_ccsd_pair_energies(self._h, _ptr(p)))
        return self._shape(p, (self.No, self.No))

    def energies(self):
        n = ctypes.c_int()
        _check(self._lib.sisi_ccsd_get_energies(self._h, None, 0, ctypes.byref(n)))
        out = np.empty(max(n.value, 1))
        _check(self._lib.sisi_ccsd_get_energies(self._h, _ptr(out), n.value, ctypes.byref(n)))
        return out[:n.value].copy()

    # --- (T) ---------------------------------------------------------------
    def triples(self, T1=None, T2=None, precision=FP64, blocks=None):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        e = ctypes.c_double()
        t1 = self._t1(T1) if T1 is not None else None
        t2 = self._t2(T2) if T2 is not None else None
        if blocks is None:
            _check(self._lib.sisi_triples_energy(self._h, _ptr(t1), _ptr(t2), int(precision),
                                                 ctypes.byref(e)))
        else:
            bl = np.ascontiguousarray(np.asarray(blocks, dtype=np.int32).reshape(-1))
            _check(self._lib.sisi_triples_energy_blocks(
                self._h, _ptr(t1), _ptr(t2), int(precision), bl.ctypes.data_as(_c_int_p),
                len(bl) // 3, ctypes.byref(e)))
        return e.value

    # --- timers ------------------------------------------------------------
    def enable_stats(self, on=True):
        _check(self._lib.sisi_engine_enable_stats(self._h, int(bool(on))))

    def reset_stats(self):
        _check(self._lib.sisi_engine_reset_stats(self._h))

    def stats(self):
        s = Stats()
        _check(self._lib.sisi_engine_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()
