"""CPU restatement of the index map of `symmetrize_doubles_kernel`
(sisi4s_amd/csrc/tensor_ops.hip): every block, (i, j) step and thread of the
launch is replayed on a host array, so that the tile-pair assignment is
checked for coverage (each element of T2 written exactly once, to the average
of itself and its partner T2[b,a,j,i]) and for bounds, including ragged tiles
(Nv not a multiple of 32) and a grid whose z extent is smaller than No*No
(the kernel then strides over the (i, j) pairs)."""

import numpy as np
import pytest


def replay(T, Nv, No, gz):
    # the engine's layout: first index fastest (Fortran order), so that the
    # flat views below alias the arrays
    T = np.array(T, order="F")
    src = T.copy(order="F")
    writes = np.zeros(T.shape, dtype=np.int64, order="F")
    tiles = (Nv + 31) // 32
    x = np.arange(256) & 31
    y0 = np.arange(256) >> 5
    flat = src.reshape(-1, order="F")
    out = T.reshape(-1, order="F")
    cnt = writes.reshape(-1, order="F")
    assert np.shares_memory(out, T) and np.shares_memory(cnt, writes)
    nv2 = Nv * Nv
    for ta in range(tiles):
        for tb in range(tiles):
            for z in range(gz):
                for ij in range(z, No * No, gz):
                    i, j = ij % No, ij // No
                    if i > j or (i == j and ta > tb):
                        continue
                    P, Q = nv2 * (i + No * j), nv2 * (j + No * i)
                    for r in range(4):
                        y = y0 + 8 * r
                        pa, pb, qa, qb = 32 * ta + x, 32 * tb + y, 32 * tb + x, 32 * ta + y
                        okp = (pa < Nv) & (pb < Nv)
                        okq = (qa < Nv) & (qb < Nv)
                        # s0[y][x] = P[pa, pb], s1[y][x] = Q[qa, qb]; partner values
                        # s1[x][y] = Q[32tb + y, 32ta + x], s0[x][y] = P[32ta + y, 32tb + x]
                        pp = P + pa + Nv * pb
                        qq = Q + qa + Nv * qb
                        partner_p = Q + (32 * tb + y) + Nv * (32 * ta + x)
                        partner_q = P + (32 * ta + y) + Nv * (32 * tb + x)
                        for k in np.nonzero(okp)[0]:
                            out[pp[k]] = 0.5 * (flat[pp[k]] + flat[partner_p[k]])
                            cnt[pp[k]] += 1
                        for k in np.nonzero(okq)[0]:
                            out[qq[k]] = 0.5 * (flat[qq[k]] + flat[partner_q[k]])
                            cnt[qq[k]] += 1
    return T, writes


@pytest.mark.parametrize("Nv,No,gz", [(5, 2, 4), (37, 3, 9), (40, 3, 4), (64, 2, 1)])
def test_symmetrize_tile_pairs_cover_every_element(Nv, No, gz):
    rng = np.random.default_rng(Nv * 100 + No)
    T = rng.standard_normal((Nv, Nv, No, No))
    out, writes = replay(T, Nv, No, gz)
    ref = 0.5 * (T + T.transpose(1, 0, 3, 2))
    # diagonal tiles of i == j write each element twice with the same value
    # (the tile is its own partner); every other element exactly once
    assert writes.min() >= 1
    a = np.arange(Nv)
    same_tile = (a[:, None] // 32) == (a[None, :] // 32)
    diag_ij = np.eye(No, dtype=bool)
    twice = same_tile[:, :, None, None] & diag_ij[None, None, :, :]
    assert np.all(writes[twice] == 2) and np.all(writes[~twice] == 1)
    assert np.array_equal(out, ref)
    assert np.array_equal(out, out.transpose(1, 0, 3, 2))
