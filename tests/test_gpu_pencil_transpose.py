"""GPU parity of every P > 1 transpose and bricks <-> pencils reshape, all ranks of the
sub-communicator run on the one GPU:
each rank's pack kernel (GPU kernel plugin, neighbour data and kernel type of
reshape_handle_generic%create), a device copy of every peer chunk in place of the
all-to-all, each rank's unpack kernel.  Each rank's receive pencil must equal the numpy
slice of the permuted global array bit for bit, for both directions of X<->Y, Y<->Z and
X<->Z (the 3D backward ones through the two-step PERMUTE_BACKWARD_START / _END pair in
modes "mpi" and "pipelined"), uneven splits, and element sizes 4, 8 and 16 bytes."""
import ctypes as C

import numpy as np
import pytest

from dtfft_amd import pencil

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")

AXES = {1: (0, 1, 2), 2: (1, 2, 0), 3: (2, 0, 1)}
AXES_2D = {1: (0, 1), 2: (1, 0)}
PENCIL_OF = {pencil.X_TO_Y: (1, 2), pencil.Y_TO_Z: (2, 3), pencil.X_TO_Z: (1, 3)}
DT = {4: np.uint32, 8: np.uint64, 16: np.dtype([("a", np.uint64), ("b", np.uint64)])}


def block(Gf, axes, counts, starts):
    sl = [slice(None)] * Gf.ndim
    for k, g in enumerate(axes):
        sl[g] = slice(starts[k], starts[k] + counts[k])
    return np.ascontiguousarray(np.transpose(Gf[tuple(sl)], axes).ravel(order="F"))


def run_all_ranks(tt, dims, P, mode, es):
    print(f"case: transpose {tt} dims {dims} P={P} mode={mode} es={es}", flush=True)
    nd = len(dims)
    n = int(np.prod(dims))
    vals = np.arange(n, dtype=np.uint64) * 2654435761 % (1 << 61)
    if es == 16:
        G = np.empty(n, dtype=DT[16])
        G["a"], G["b"] = vals, vals ^ 0x5555
    else:
        G = vals.astype(DT[es])
    Gf = G.reshape(dims[::-1]).transpose(*range(nd)[::-1])
    a, b = PENCIL_OF[abs(tt)]
    pin, pout = (a, b) if tt > 0 else (b, a)
    ax = AXES_2D if nd == 2 else AXES
    ranks = [D.PencilTranspose(tt, dims, P, r, es, mode) for r in range(P)]
    ins, sends, recvs, outs = [], [], [], []
    for t in ranks:
        sh = t.share
        x = block(Gf, ax[pin], sh.in_counts, sh.in_starts)
        ins.append(D.DeviceArray.from_numpy(x.view(np.uint8), 16))
        sends.append(D.DeviceArray(max(16, t.send_bytes)))
        recvs.append(D.DeviceArray(max(16, t.recv_bytes)))
        outs.append(D.DeviceArray(max(16, int(np.prod(sh.out_counts)) * es)))
    for t, i_, s_ in zip(ranks, ins, sends):
        t.pack_into(i_, s_)
    for r, t in enumerate(ranks):
        for i, u in enumerate(ranks):
            nbytes = u.share.send_counts[r] * es
            if nbytes:
                D.call("dtfft_hip_memcpy_d2d_async", C.c_void_p(recvs[r].ptr + t.share.recv_displs[i] * es),
                       C.c_void_p(sends[i].ptr + u.share.send_displs[r] * es), C.c_size_t(nbytes), None)
    for t, r_, o_ in zip(ranks, recvs, outs):
        t.unpack_from(r_, o_)
    D.synchronize()
    for r, t in enumerate(ranks):
        sh = t.share
        want = block(Gf, ax[pout], sh.out_counts, sh.out_starts)
        got = outs[r].download(want.size * es, np.uint8)
        assert got.tobytes() == want.view(np.uint8).tobytes(), (tt, dims, P, mode, es, r)


@pytest.mark.parametrize("es", [4, 8, 16])
@pytest.mark.parametrize("mode", pencil.MODES)
@pytest.mark.parametrize("tt", [pencil.X_TO_Y, pencil.Y_TO_X, pencil.Y_TO_Z, pencil.Z_TO_Y,
                                pencil.X_TO_Z, pencil.Z_TO_X])
@pytest.mark.parametrize("dims,P", [((17, 10, 7), 3), ((33, 77, 21), 4), ((5, 3, 11), 4)])
def test_every_transpose_3d_all_ranks(tt, dims, P, mode, es):
    run_all_ranks(tt, dims, P, mode, es)


@pytest.mark.parametrize("es", [8, 16])
@pytest.mark.parametrize("mode", pencil.MODES)
@pytest.mark.parametrize("tt", [pencil.X_TO_Y, pencil.Y_TO_X])
@pytest.mark.parametrize("dims,P", [((18, 155), 4), ((90, 57), 3)])
def test_every_transpose_2d_all_ranks(tt, dims, P, mode, es):
    run_all_ranks(tt, dims, P, mode, es)


@pytest.mark.parametrize("mode", ["mpi", "fused"])
@pytest.mark.parametrize("tt", [pencil.Z_TO_Y, pencil.Y_TO_X])
def test_backward_share_of_2048_grid(tt, mode):
    """The backward direction at the benchmark share's scale: 4 ranks of a 2048^3 complex64
    grid's Z->Y (or Y->X) exchange on local slabs of 64 (reduced so that all four ranks fit
    the one GPU's test time), uneven by one point on the exchanged axes."""
    run_all_ranks(tt, (258, 2047, 131) if tt == pencil.Z_TO_Y else (2047, 258, 33), 4, mode, 8)


# ---- bricks <-> pencils reshapes, all ranks on the one GPU --------------------------------
RESHAPES = [pencil.X_BRICKS_TO_PENCILS, pencil.X_PENCILS_TO_BRICKS, pencil.Z_PENCILS_TO_BRICKS,
            pencil.Z_BRICKS_TO_PENCILS]


def run_reshape_all_ranks(rt, dims, P, mode, es, split=None):
    print(f"case: reshape {rt} dims {dims} P={P} split={split} mode={mode} es={es}", flush=True)
    nd = len(dims)
    n = int(np.prod(dims))
    vals = np.arange(n, dtype=np.uint64) * 2654435761 % (1 << 61)
    if es == 16:
        G = np.empty(n, dtype=DT[16])
        G["a"], G["b"] = vals, vals ^ 0x5555
    else:
        G = vals.astype(DT[es])
    Gf = G.reshape(dims[::-1]).transpose(*range(nd)[::-1])
    ax = tuple(range(nd))
    ranks = [D.PencilReshape(rt, dims, P, r, es, mode, split) for r in range(P)]
    ins, sends, recvs, outs = [], [], [], []
    for t in ranks:
        sh = t.share
        x = block(Gf, ax, sh.in_counts, sh.in_starts)
        ins.append(D.DeviceArray.from_numpy(x.view(np.uint8), 16))
        # pack-free: the input is the send buffer
        sends.append(ins[-1] if t.pack is None else D.DeviceArray(max(16, t.send_bytes)))
        recvs.append(D.DeviceArray(max(16, t.recv_bytes)))
        # unpack-free: the receive buffer is the output
        outs.append(recvs[-1] if t.unpack is None else D.DeviceArray(max(16, int(np.prod(sh.out_counts)) * es)))
    for t, i_, s_ in zip(ranks, ins, sends):
        t.pack_into(i_, s_)
    for r, t in enumerate(ranks):
        for i, u in enumerate(ranks):
            nbytes = u.share.send_counts[r] * es
            if nbytes:
                D.call("dtfft_hip_memcpy_d2d_async", C.c_void_p(recvs[r].ptr + t.share.recv_displs[i] * es),
                       C.c_void_p(sends[i].ptr + u.share.send_displs[r] * es), C.c_size_t(nbytes), None)
    for t, r_, o_ in zip(ranks, recvs, outs):
        t.unpack_from(r_, o_)
    D.synchronize()
    for r, t in enumerate(ranks):
        sh = t.share
        want = block(Gf, ax, sh.out_counts, sh.out_starts)
        got = outs[r].download(want.size * es, np.uint8)
        assert got.tobytes() == want.view(np.uint8).tobytes(), (rt, dims, P, mode, es, r, sh.strat)


@pytest.mark.parametrize("es", [4, 8, 16])
@pytest.mark.parametrize("mode", pencil.RESHAPE_MODES)
@pytest.mark.parametrize("rt", RESHAPES)
@pytest.mark.parametrize("dims,P,split", [((17, 10, 7), 3, (1, 3)), ((9, 23, 13), 4, (4, 1)),
                                          ((12, 11, 9), 4, (2, 2)), ((5, 3, 11), 4, (2, 2))])
def test_reshape_3d_all_ranks(rt, dims, P, split, mode, es):
    run_reshape_all_ranks(rt, dims, P, mode, es, split)


@pytest.mark.parametrize("mode", pencil.RESHAPE_MODES)
@pytest.mark.parametrize("rt", RESHAPES)
@pytest.mark.parametrize("dims,P", [((18, 155), 4), ((90, 57), 3)])
def test_reshape_2d_all_ranks(rt, dims, P, mode):
    run_reshape_all_ranks(rt, dims, P, mode, 8)
