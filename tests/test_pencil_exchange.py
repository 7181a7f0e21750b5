"""CPU check of the P > 1 transpose layout (dtfft_amd.pencil) by simulating all ranks
of a sub-communicator with the oracle's host kernels: every rank packs with
KERNEL_PACK_FORWARD, the chunks are exchanged all-to-all in memory, every rank unpacks
with KERNEL_UNPACK -- and the result must equal the global axis permutation computed
by numpy (independent of both the kernels and the neighbour data)."""
import numpy as np
import pytest

import oracle as O
from dtfft_amd import pencil


def simulate(a, b, c, P, dtype=np.uint64):
    G = np.arange(a * b * c, dtype=dtype).reshape(c, b, a)  # Fortran [a, b, c]
    bparts = pencil.split(b, P)
    aparts = pencil.split(a, P)
    layouts = [pencil.forward_exchange(a, b, c, P, r) for r in range(P)]
    sends = []
    for r in range(P):
        b0, bn = bparts[r]
        A = np.ascontiguousarray(G[:, b0:b0 + bn, :]).ravel()
        srows, _, scount, _ = layouts[r]
        buf = np.zeros(sum(scount), dtype=dtype)
        for i in range(P):
            O.host_kernel(O.KERNEL_PACK_FORWARD, (a, bn, c), A, buf, srows, i + 1)
        sends.append(buf)
    results = []
    for r in range(P):
        _, rrows, _, rcount = layouts[r]
        recv = np.zeros(sum(rcount), dtype=dtype)
        for i in range(P):
            srows_i = layouts[i][0]
            off, n = srows_i[r][4], layouts[i][2][r]
            assert n == rcount[i]
            recv[rrows[i][3]:rrows[i][3] + n] = sends[i][off:off + n]
        a0, an = aparts[r]
        B = np.zeros(b * c * an, dtype=dtype)
        O.host_kernel(O.KERNEL_UNPACK, (b, c, an), recv, B, rrows)
        want = np.transpose(G[:, :, a0:a0 + an], (2, 0, 1)).ravel()
        results.append((B, want))
    return results


@pytest.mark.parametrize("a,b,c,P", [(16, 16, 4, 4), (17, 10, 3, 3), (9, 23, 5, 4), (2048 // 32, 512 // 32, 2, 4)])
def test_forward_exchange_equals_global_permutation(a, b, c, P):
    for B, want in simulate(a, b, c, P):
        assert np.array_equal(B, want)


def test_local_size_remainder_to_last_ranks():
    assert pencil.split(10, 4) == [(0, 2), (2, 2), (4, 3), (7, 3)]
    assert pencil.split(8, 1) == [(0, 8)]


def test_config5_rows():
    """One-GPU share of a 4x4 decomposition of 2048^3: Y-pencil [2048, 512, 512] -> Z."""
    send, recv, sc, rc = pencil.forward_exchange(2048, 2048, 512, 4, 1)
    assert send[2] == (512, 512, 512, 1024, 2 * 512 ** 3)
    assert recv[3] == (512, 512, 512, 3 * 512 ** 3, 1536)
    assert sc == rc == [512 ** 3] * 4
