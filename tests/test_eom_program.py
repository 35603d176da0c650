"""The contraction programs of the device EOM right-applies (eom_ip_sigma and
eom_ea_sigma in sisi4s_amd/csrc/eom.hip, after the block builds they call:
make_taus, build_common_blocks and build_ip_blocks / build_ea_blocks)
replayed on the CPU: each ctr / sum /
copy / zero / add_t1t1 / add_diagonal statement of the function body runs as
the numpy operation the engine's planner documents (ctr: C = alpha einsum(A, B)
+ beta C; sum: C[ic] = alpha A[ia] + beta C[ic]; first index fastest), on
buffers that start as NaN so a read before a write shows, and the resulting
sigma vectors are compared with the oracle's.  This checks the index
strings, signs, factors and buffer reuse of the device code without a GPU;
the IP program, which the GPU tests check against the oracle, also checks
the replay itself.  Any statement the replay does not know fails the test."""

import os
import re

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "sisi4s_amd", "csrc", "eom.hip")


def _body(name):
    text = open(SRC).read()
    start = text.index(" %s(Engine &e" % name)
    i = text.index("{", start)
    depth, j = 0, i
    while True:
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
            if depth == 0:
                break
        j += 1
    body = text[i + 1:j]
    return re.sub(r"//[^\n]*", "", body)


def _split_top(s):
    """Split on commas outside parentheses, braces, template brackets and
    string literals."""
    out, depth, cur, quoted = [], 0, "", False
    for ch in s:
        if ch == '"':
            quoted = not quoted
        elif not quoted and ch in "({<":
            depth += 1
        elif not quoted and ch in ")}>":
            depth -= 1
        if ch == "," and depth == 0 and not quoted:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


class Replay:
    def __init__(self, o, v, buffers):
        self.env = dict(o=o, v=v, n2=v * v * o * o)
        self.mem = {k: np.asarray(x, dtype=float).ravel(order="F").copy() for k, x in buffers.items()}
        self.alias, self.views = {}, {}

    def key(self, expr):
        expr = expr.strip()
        m = re.fullmatch(r"const_cast<double \*>\((.*)\)", expr)
        if m:
            expr = m.group(1).strip()
        return self.alias.get(expr, expr)

    def num(self, expr):
        expr = re.sub(r"int\((\w+)\)", r"\1", expr)
        return eval(expr, {}, dict(self.env))

    def get(self, name):
        k, dims = self.views[name]
        n = int(np.prod(dims))
        flat = self.mem.setdefault(k, np.full(n, np.nan))
        assert flat.size >= n, (name, k)
        return flat[:n].reshape(dims, order="F")

    def put(self, name, x):
        k, dims = self.views[name]
        n = int(np.prod(dims))
        flat = self.mem.setdefault(k, np.full(n, np.nan))
        if flat.size < n:
            flat = self.mem[k] = np.concatenate([flat, np.full(n - flat.size, np.nan)])
        assert x.shape == tuple(dims), (name, x.shape, dims)
        flat[:n] = x.ravel(order="F")

    def flat(self, expr, n):
        k = self.key(expr)
        flat = self.mem.setdefault(k, np.full(n, np.nan))
        if flat.size < n:
            flat = self.mem[k] = np.concatenate([flat, np.full(n - flat.size, np.nan)])
        return flat

    def run(self, body):
        for st in body.split(";"):
            st = " ".join(st.split())
            if not st:
                continue
            if st.startswith(("SISI_REQUIRE", "e.thaw()", "const int64_t", "const int p", "DBuf ",
                              "if (", "return", "Taus ", "build_", "e.eom.ip = ", "e.eom.ea = ")):
                continue
            if re.fullmatch(r"e\.eom\.\w+ = DBuf\(.*\)", st):
                continue  # a block's allocation
            if st.startswith("double *"):
                for m in re.finditer(r"\*(\w+) = const_cast<double \*>\((\w+)\)", st):
                    self.alias[m.group(1)] = m.group(2)
                continue
            if st.startswith("View ") or re.fullmatch(r"[\w.]+ = dense\(.*\)", st):
                for d in _split_top(st[5:] if st.startswith("View ") else st):
                    m = re.fullmatch(r"([\w.]+) = dense\((.*), \{(.*)\}\)", d)
                    assert m, d
                    dims = [self.num(x) for x in m.group(3).split(",")]
                    self.views[m.group(1)] = (self.key(m.group(2)), dims)
                continue
            m = re.fullmatch(r"(\w+)\((.*)\)", st)
            assert m, "unknown statement: " + st
            fn, args = m.group(1), _split_top(m.group(2))
            assert args[0] == "e", st
            if fn == "ctr":
                alpha, spec, A, B, beta, C = float(args[1]), args[2].strip('"'), args[3], args[4], \
                    float(args[5]), args[6]
                r = alpha * np.einsum(spec, self.get(A), self.get(B))
                self.put(C, r + beta * self.get(C) if beta != 0.0 else r)
            elif fn == "sum":
                alpha, A, ia, beta, C, ic = float(args[1]), args[2], args[3].strip('"'), \
                    float(args[4]), args[5], args[6].strip('"')
                r = alpha * np.einsum(ia + "->" + ic, self.get(A))
                self.put(C, r + beta * self.get(C) if beta != 0.0 else r)
            elif fn == "copy":
                dst, src = args[1], args[2]
                alpha = float(args[3]) if len(args) > 3 else 1.0
                beta = float(args[4]) if len(args) > 4 else 0.0
                r = alpha * self.get(src)
                self.put(dst, r + beta * self.get(dst) if beta != 0.0 else r)
            elif fn == "zero":
                n = int(self.num(args[2]))
                self.flat(args[1], n)[:n] = 0.0
            elif fn == "add_t1t1":
                X, T2, T1, f = args[1], args[2], args[3], float(args[4])
                v, o = int(self.num(args[5])), int(self.num(args[6]))
                n = v * v * o * o
                t1 = self.flat(T1, v * o)[:v * o].reshape((v, o), order="F")
                t2 = self.flat(T2, n)[:n].reshape((v, v, o, o), order="F")
                x = t2 + f * np.einsum("ai,bj->abij", t1, t1)
                self.flat(X, n)[:n] = x.ravel(order="F")
            elif fn == "add_diagonal":
                n = int(self.num(args[2]))
                M = self.flat(args[1], n * n)
                M[:n * n:n + 1] += self.flat(args[3], n)[:n]
            else:
                raise AssertionError("unknown statement: " + st)


def _system(no, n, seed):
    from oracle import uccsd
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(12, n, n)) * 0.25
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    asym = eri - np.swapaxes(eri, 2, 3)
    eps = np.concatenate([np.linspace(-1.3, -0.8, no), np.sort(rng.uniform(0.3, 2.0, n - no))])
    ints = uccsd.classes(asym, no)
    # generic amplitudes: the right-apply is linear in r for any T
    nv = n - no
    T1 = 0.1 * rng.normal(size=(nv, no))
    T2 = 0.1 * rng.normal(size=(nv, nv, no, no))
    T2 = T2 - T2.transpose(1, 0, 2, 3)
    T2 = T2 - T2.transpose(0, 1, 3, 2)
    return ints, eps[:no], eps[no:], T1, T2, rng


def _buffers(ints, ei, ea, T1, T2):
    return {"e.eom.T1.get()": T1, "e.eom.T2.get()": T2, "e.Vabij.get()": ints["PPHH"], "e.Vijkl.get()": ints["HHHH"],
            "e.Vijka.get()": ints["HHHP"], "e.Vabci.get()": ints["PPPH"], "e.Vaibj.get()": ints["PHPH"],
            "e.Vabcd.get()": ints["PPPP"], "e.eps_i.get()": ei, "e.eps_a.get()": ea}


@pytest.mark.parametrize("no,n", [(3, 8), (4, 9)])
def test_ip_program_replays_to_the_oracle(no, n):
    from oracle import eom
    ints, ei, ea, T1, T2, rng = _system(no, n, 21)
    nv = n - no
    r1 = rng.normal(size=no)
    r2 = rng.normal(size=(no, no, nv))
    r2 = r2 - r2.transpose(1, 0, 2)
    rp = Replay(no, nv, dict(_buffers(ints, ei, ea, T1, T2), r1c=r1, r2c=r2))
    for f in ("make_taus", "build_common_blocks", "build_ip_blocks", "eom_ip_sigma"):
        rp.run(_body(f))
    s1 = rp.mem["s1p"][:no]
    s2 = rp.mem["s2p"][:no * no * nv].reshape((no, no, nv), order="F")
    H = eom.hbar_ip(ints, ei, ea, T1, T2)
    w1, w2 = eom.ip_sigma(H, r1, r2)
    assert np.abs(s1 - w1).max() <= 1e-12 * max(1.0, np.abs(w1).max())
    assert np.abs(s2 - w2).max() <= 1e-12 * max(1.0, np.abs(w2).max())


@pytest.mark.parametrize("no,n", [(3, 8), (4, 9)])
def test_ea_program_replays_to_the_oracle(no, n):
    from oracle import eom
    ints, ei, ea, T1, T2, rng = _system(no, n, 22)
    nv = n - no
    r1 = rng.normal(size=nv)
    r2 = rng.normal(size=(no, nv, nv))
    r2 = r2 - r2.transpose(0, 2, 1)
    rp = Replay(no, nv, dict(_buffers(ints, ei, ea, T1, T2), r1c=r1, r2c=r2))
    for f in ("make_taus", "build_common_blocks", "build_ea_blocks", "eom_ea_sigma"):
        rp.run(_body(f))
    s1 = rp.mem["s1p"][:nv]
    s2 = rp.mem["s2p"][:no * nv * nv].reshape((no, nv, nv), order="F")
    H = eom.hbar_ea(ints, ei, ea, T1, T2)
    w1, w2 = eom.ea_sigma(H, r1, r2)
    assert np.abs(s1 - w1).max() <= 1e-12 * max(1.0, np.abs(w1).max())
    assert np.abs(s2 - w2).max() <= 1e-12 * max(1.0, np.abs(w2).max())


def test_replay_catches_a_planted_sign_error():
    """The replay is sensitive: the IP program with one sign flipped in its
    text (a ring term of W_mbij) no longer matches the oracle."""
    from oracle import eom
    ints, ei, ea, T1, T2, rng = _system(3, 8, 23)
    r1, r2 = rng.normal(size=3), rng.normal(size=(3, 3, 5))
    r2 = r2 - r2.transpose(1, 0, 2)
    body = _body("build_ip_blocks")
    planted = body.replace('ctr(e, 1.0, "mnie,bejn->mbij"', 'ctr(e, -1.0, "mnie,bejn->mbij"')
    assert planted != body
    rp = Replay(3, 5, dict(_buffers(ints, ei, ea, T1, T2), r1c=r1, r2c=r2))
    for f in ("make_taus", "build_common_blocks"):
        rp.run(_body(f))
    rp.run(planted)
    rp.run(_body("eom_ip_sigma"))
    s2 = rp.mem["s2p"][:45].reshape((3, 3, 5), order="F")
    _, w2 = eom.ip_sigma(eom.hbar_ip(ints, ei, ea, T1, T2), r1, r2)
    assert np.abs(s2 - w2).max() > 1e-4
