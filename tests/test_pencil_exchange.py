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


# ---- every transpose type, both directions, every exchange mode, uneven splits ----------
AXES = {1: (0, 1, 2), 2: (1, 2, 0), 3: (2, 0, 1)}  # X, Y, Z pencils: global axis of each local axis
AXES_2D = {1: (0, 1), 2: (1, 0)}
PENCIL_OF = {pencil.X_TO_Y: (1, 2), pencil.Y_TO_Z: (2, 3), pencil.X_TO_Z: (1, 3)}


def pencils_of(tt, ndims):
    a, b = PENCIL_OF[abs(tt)]
    return (a, b) if tt > 0 else (b, a)


def block(Gf, axes, counts, starts):
    """Local block of the Fortran-indexed global array Gf in a pencil's axis order, raveled
    in Fortran order."""
    sl = [slice(None)] * Gf.ndim
    for k, g in enumerate(axes):
        sl[g] = slice(starts[k], starts[k] + counts[k])
    return np.ascontiguousarray(np.transpose(Gf[tuple(sl)], axes).ravel(order="F"))


def simulate_all(tt, dims, P, mode):
    nd = len(dims)
    Gf = np.arange(int(np.prod(dims)), dtype=np.uint64).reshape(dims[::-1]).transpose(*range(nd)[::-1])
    pin, pout = pencils_of(tt, nd)
    ax = AXES_2D if nd == 2 else AXES
    shares = [pencil.TransposeShare(tt, dims, P, r, mode) for r in range(P)]
    sends = []
    for s in shares:
        inbuf = block(Gf, ax[pin], s.in_counts, s.in_starts)
        buf = np.zeros(max(1, sum(s.send_counts)), dtype=np.uint64)
        rows = np.array(s.send_rows)
        s.run_pack(lambda i, s=s, inbuf=inbuf, buf=buf, rows=rows: O.host_kernel(
            s.pack_kernel, s.in_counts, inbuf, buf, rows, i if i else None))
        sends.append(buf)
    for r, s in enumerate(shares):
        recv = np.zeros(max(1, sum(s.recv_counts)), dtype=np.uint64)
        for i, t in enumerate(shares):
            n = t.send_counts[r]
            assert n == s.recv_counts[i]
            recv[s.recv_displs[i]:s.recv_displs[i] + n] = sends[i][t.send_displs[r]:t.send_displs[r] + n]
        out = np.zeros(int(np.prod(s.out_counts)), dtype=np.uint64)
        rows = np.array(s.recv_rows)
        s.run_unpack(lambda i, s=s, recv=recv, out=out, rows=rows: O.host_kernel(
            s.unpack_kernel, s.out_counts, recv, out, rows, i if i else None))
        want = block(Gf, ax[pout], s.out_counts, s.out_starts)
        assert np.array_equal(out, want), (tt, dims, P, mode, r)


@pytest.mark.parametrize("mode", pencil.MODES)
@pytest.mark.parametrize("tt", [pencil.X_TO_Y, pencil.Y_TO_X, pencil.Y_TO_Z, pencil.Z_TO_Y,
                                pencil.X_TO_Z, pencil.Z_TO_X])
@pytest.mark.parametrize("dims,P", [((17, 10, 7), 3), ((9, 23, 13), 4), ((8, 8, 8), 2), ((5, 3, 11), 4)])
def test_every_transpose_3d_equals_global_permutation(tt, dims, P, mode):
    """All ranks of the sub-communicator simulated with the oracle's host kernels: neighbour
    data, kernel types and the pack -> all-to-all -> unpack order must rebuild each rank's
    receive pencil of the global array (numpy slices, independent of both).  Uneven splits
    put the remainder on the last ranks; (5, 3, 11) on 4 ranks leaves one rank without Y."""
    simulate_all(tt, dims, P, mode)


@pytest.mark.parametrize("mode", pencil.MODES)
@pytest.mark.parametrize("tt", [pencil.X_TO_Y, pencil.Y_TO_X])
@pytest.mark.parametrize("dims,P", [((18, 155), 4), ((90, 57), 3), ((7, 5), 2)])
def test_every_transpose_2d_equals_global_permutation(tt, dims, P, mode):
    simulate_all(tt, dims, P, mode)


def test_kernel_selection():
    """Kernel types of src/dtfft_reshape_handle_generic.F90:401-413, 582-600."""
    kt = pencil.transpose_kernel_types
    assert kt(pencil.X_TO_Y, 3, "mpi") == (pencil.KERNEL_PERMUTE_FORWARD, pencil.KERNEL_UNPACK, False)
    assert kt(pencil.Z_TO_Y, 3, "mpi") == (pencil.KERNEL_PERMUTE_BACKWARD_START,
                                           pencil.KERNEL_PERMUTE_BACKWARD_END, True)
    assert kt(pencil.Y_TO_X, 3, "pipelined") == (pencil.KERNEL_PERMUTE_BACKWARD_START,
                                                 pencil.KERNEL_PERMUTE_BACKWARD_END_PIPELINED, True)
    assert kt(pencil.Y_TO_X, 3, "fused") == (pencil.KERNEL_PACK_BACKWARD, pencil.KERNEL_UNPACK_PIPELINED, False)
    assert kt(pencil.Y_TO_X, 2, "mpi") == (pencil.KERNEL_PERMUTE_BACKWARD, pencil.KERNEL_UNPACK, False)
    assert kt(pencil.X_TO_Z, 3, "mpi") == (pencil.KERNEL_PERMUTE_BACKWARD, pencil.KERNEL_UNPACK, False)
    assert kt(pencil.Z_TO_X, 3, "unpack") == (pencil.KERNEL_PACK, pencil.KERNEL_UNPACK_FORWARD, False)
    assert kt(pencil.Z_TO_Y, 3, "unpack") == (pencil.KERNEL_PACK, pencil.KERNEL_UNPACK_BACKWARD, True)


def test_forward_rows_match_general_rows():
    """The equal-split share of the benchmark (forward_exchange) is a case of the general
    neighbour data: Y->Z of a 2048 x 2048 x 2048 grid, rank 1 of 4, local X count 512."""
    send, recv, sc, rc = pencil.forward_exchange(2048, 2048, 512, 4, 1)
    s = pencil.TransposeShare(pencil.Y_TO_Z, (512 * 4, 2048, 2048), 4, 1, "fused", fixed=(512, 0))
    assert [tuple(r) for r in s.send_rows] == send and [tuple(r) for r in s.recv_rows] == recv
    assert s.send_counts == sc and s.recv_counts == rc


# ---- bricks <-> pencils reshapes ----------------------------------------------------------
def simulate_reshape(rt, dims, P, mode, split=None):
    nd = len(dims)
    Gf = np.arange(int(np.prod(dims)), dtype=np.uint64).reshape(dims[::-1]).transpose(*range(nd)[::-1])
    ax = tuple(range(nd))  # both layouts in the same axis order
    shares = [pencil.ReshapeShare(rt, dims, P, r, mode, split) for r in range(P)]
    sends = []
    for s in shares:
        inbuf = block(Gf, ax, s.in_counts, s.in_starts)
        if s.pack_kernel == pencil.KERNEL_DUMMY:
            sends.append(inbuf)  # pack-free: the chunks already lie at their displacements
            continue
        buf = np.zeros(max(1, sum(s.send_counts)), dtype=np.uint64)
        rows = np.array(s.send_rows)
        s.run_pack(lambda i, s=s, inbuf=inbuf, buf=buf, rows=rows: O.host_kernel(
            s.pack_kernel, s.in_counts, inbuf, buf, rows, i if i else None))
        sends.append(buf)
    for r, s in enumerate(shares):
        recv = np.zeros(max(1, sum(s.recv_counts)), dtype=np.uint64)
        for i, t in enumerate(shares):
            n = t.send_counts[r]
            assert n == s.recv_counts[i]
            recv[s.recv_displs[i]:s.recv_displs[i] + n] = sends[i][t.send_displs[r]:t.send_displs[r] + n]
        want = block(Gf, ax, s.out_counts, s.out_starts)
        if s.unpack_kernel == pencil.KERNEL_DUMMY:
            out = recv[:want.size]  # unpack-free: the chunks land in the final layout
        else:
            out = np.zeros(int(np.prod(s.out_counts)), dtype=np.uint64)
            rows = np.array(s.recv_rows)
            s.run_unpack(lambda i, s=s, recv=recv, out=out, rows=rows: O.host_kernel(
                s.unpack_kernel, s.out_counts, recv, out, rows, i if i else None))
        assert np.array_equal(out, want), (rt, dims, P, mode, split, r, s.strat)
    return shares


RESHAPES = [pencil.X_BRICKS_TO_PENCILS, pencil.X_PENCILS_TO_BRICKS, pencil.Z_PENCILS_TO_BRICKS,
            pencil.Z_BRICKS_TO_PENCILS]


@pytest.mark.parametrize("mode", pencil.RESHAPE_MODES)
@pytest.mark.parametrize("rt", RESHAPES)
@pytest.mark.parametrize("dims,P,split,strat", [
    ((17, 10, 7), 3, (1, 3), 1),    # pencils split along the third axis only: pack/unpack-free
    ((9, 23, 13), 4, (4, 1), 2),    # along the second axis only
    ((12, 11, 9), 4, (2, 2), 3),    # along both
    ((8, 8, 8), 6, (3, 2), 3),
    ((5, 3, 11), 4, (2, 2), 3),     # a rank with nothing on one axis
])
def test_reshape_3d_equals_global_layout(rt, dims, P, split, strat, mode):
    """All ranks of a reshape group simulated with the oracle's host kernels (pack, all-to-all,
    unpack, or none of the copies where the strategy makes them unnecessary) must rebuild each
    rank's target layout of the global array."""
    shares = simulate_reshape(rt, dims, P, mode, split)
    assert all(s.strat == strat for s in shares)
    b2p = rt in pencil.BRICKS_TO_PENCILS
    assert all(s.pack_free == (b2p and strat == 1) for s in shares)
    assert all(s.unpack_free == ((not b2p) and strat == 1) for s in shares)


@pytest.mark.parametrize("mode", pencil.RESHAPE_MODES)
@pytest.mark.parametrize("rt", RESHAPES)
@pytest.mark.parametrize("dims,P", [((18, 155), 4), ((90, 57), 3)])
def test_reshape_2d_equals_global_layout(rt, dims, P, mode):
    simulate_reshape(rt, dims, P, mode)
