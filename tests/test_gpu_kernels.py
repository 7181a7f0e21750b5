"""GPU parity of the kernel plugin: every kernel type, element sizes 4/8/16, the
reference's test shapes, bit-exact against the oracle's restatement of
_dtfft_kernel_host_routines.inc (the same check as src/tests/test_device_kernels.F90,
device vs host, extended to multi-peer neighbour data and nonzero offsets)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")
KT = D.KernelType

SHAPES3 = [(18, 33, 155), (33, 77, 21), (5, 1, 7), (64, 64, 64)]
SHAPES2 = [(18, 155), (90, 57), (1, 9)]


def payload(n, es, seed=0):
    rng = np.random.default_rng(seed)
    raw = rng.integers(0, 2**63, size=(n, es // 4), dtype=np.uint64).astype(np.uint32)
    return raw.view(O.elem_dtype(es)).reshape(n) if es != 16 else raw.view(O.elem_dtype(16)).reshape(n)


def run_both(kt, dims, es, nd=None, neighbor=0, out_size=None, inp=None):
    n = int(np.prod(dims))
    inp = payload(n, es) if inp is None else inp
    out_size = out_size or n
    want = np.zeros(out_size, dtype=inp.dtype)
    O.host_kernel(kt, dims, inp, want, nd, neighbor or None)
    di = D.DeviceArray.from_numpy(inp)
    do = D.DeviceArray.from_numpy(np.zeros(out_size, dtype=inp.dtype))
    k = D.Kernel(kt, dims, es, nd)
    k.execute(di, do, neighbor=neighbor, sync=True)
    got = do.download(out_size, inp.dtype)
    return got, want


def same(a, b):
    return a.tobytes() == b.tobytes()


@pytest.mark.parametrize("es", [4, 8, 16])
@pytest.mark.parametrize("dims", SHAPES3)
@pytest.mark.parametrize("kt", [KT.PERMUTE_FORWARD, KT.PERMUTE_BACKWARD, KT.PERMUTE_BACKWARD_START])
def test_permute_3d(kt, dims, es):
    got, want = run_both(kt, dims, es)
    assert same(got, want)


@pytest.mark.parametrize("es", [4, 8, 16])
@pytest.mark.parametrize("dims", SHAPES2)
@pytest.mark.parametrize("kt", [KT.PERMUTE_FORWARD, KT.PERMUTE_BACKWARD])
def test_permute_2d(kt, dims, es):
    got, want = run_both(kt, dims, es)
    assert same(got, want)


def x_split(nx, parts):
    """Split [0, nx) into `parts` pieces, remainder to the last ones (src/dtfft_pencil.F90:263-267)."""
    base, res = divmod(nx, parts)
    sizes = [base + (1 if i >= parts - res else 0) for i in range(parts)]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).tolist()
    return list(zip(starts, sizes))


@pytest.mark.parametrize("es", [4, 8, 16])
@pytest.mark.parametrize("dims", [(33, 77, 21), (18, 33, 155), (90, 57)])
def test_pack_unpack_multi_peer(dims, es):
    nd3 = list(dims) + [1] * (3 - len(dims))
    nx, ny, nz = nd3
    parts = x_split(nx, 3)
    pack_nd, off = [], 0
    for x0, w in parts:
        pack_nd.append([w, ny, nz, x0, off])
        off += w * ny * nz
    # PACK (all peers) and PACK_PIPELINED (one peer at a time)
    got, want = run_both(KT.PACK, dims, es, pack_nd)
    assert same(got, want)
    for peer in (1, 2, 3):
        got, want = run_both(KT.PACK_PIPELINED, dims, es, pack_nd, neighbor=peer)
        assert same(got, want)
    unpack_nd = [[w, ny, nz, row[4], x0] for (x0, w), row in zip(parts, pack_nd)]
    got, want = run_both(KT.UNPACK, dims, es, unpack_nd)
    assert same(got, want)
    got, want = run_both(KT.UNPACK_PIPELINED, dims, es, unpack_nd, neighbor=2)
    assert same(got, want)
    # pack_forward per peer: permuted chunk of each x-slab
    for peer in (1, 3):
        got, want = run_both(KT.PACK_FORWARD, dims, es, pack_nd, neighbor=peer)
        assert same(got, want)


@pytest.mark.parametrize("es", [4, 8, 16])
@pytest.mark.parametrize("dims", [(33, 77, 21), (18, 33, 155)])
def test_pack_backward_and_two_step(dims, es):
    nx, ny, nz = dims
    parts = x_split(nx, 2)
    nd, off = [], 0
    for x0, w in parts:
        nd.append([w, ny, nz, x0, off])
        off += w * ny * nz
    for peer in (1, 2):
        got, want = run_both(KT.PACK_BACKWARD, dims, es, nd, neighbor=peer)
        assert same(got, want)
    # two-step backward on one rank: backward_end(backward_start(x)) == backward(x)
    inp = payload(nx * ny * nz, es, seed=3)
    gold = np.zeros_like(inp)
    O.host_kernel(KT.PERMUTE_BACKWARD, dims, inp, gold)
    tmp_got, _ = run_both(KT.PERMUTE_BACKWARD_START, dims, es, inp=inp)
    tdims = (nz, nx, ny)
    bnd = [[nz, nx, ny, 0, 0]]
    got, want = run_both(KT.PERMUTE_BACKWARD_END, tdims, es, bnd, inp=tmp_got)
    assert same(got, want) and same(got, gold)
    got, _ = run_both(KT.PERMUTE_BACKWARD_END_PIPELINED, tdims, es, bnd, neighbor=1, inp=tmp_got)
    assert same(got, gold)


@pytest.mark.parametrize("es", [4, 8, 16])
@pytest.mark.parametrize("dims", [(33, 77, 21), (90, 57)])
def test_unpack_forward_backward(dims, es):
    """Host-only kinds in the reference (unpack mode), supported on the GPU here."""
    udims = (dims[1], dims[0]) if len(dims) == 2 else (dims[1], dims[2], dims[0])
    nd = [list(udims) + [1] * (3 - len(dims)) + [0, 0]]
    got, want = run_both(KT.UNPACK_FORWARD, udims, es, nd)
    assert same(got, want)
    got, want = run_both(KT.UNPACK_FORWARD_PIPELINED, udims, es, nd, neighbor=1)
    assert same(got, want)
    if len(dims) == 3:
        bdims = (dims[2], dims[0], dims[1])
        bnd = [list(bdims) + [0, 0]]
        got, want = run_both(KT.UNPACK_BACKWARD, bdims, es, bnd)
        assert same(got, want)


@pytest.mark.parametrize("es", [8, 16])
def test_copy_kernels(es):
    got, want = run_both(KT.COPY, (10, 7, 3), es)
    assert same(got, want)
    nd = [[5, 2, 1, 10, 30], [4, 1, 1, 0, 0]]
    got, want = run_both(KT.COPY_PIPELINED, (10, 7, 3), es, nd, neighbor=1)
    assert same(got, want)


def test_identity_cycle_gpu():
    """forward^3 == identity on the GPU alone (no oracle involved)."""
    dims = (48, 40, 24)
    inp = payload(int(np.prod(dims)), 16, seed=9)
    a = D.DeviceArray.from_numpy(inp)
    b = D.DeviceArray.from_numpy(np.zeros_like(inp))
    k1 = D.Kernel(KT.PERMUTE_FORWARD, dims, 16)
    k2 = D.Kernel(KT.PERMUTE_FORWARD, (dims[1], dims[2], dims[0]), 16)
    k3 = D.Kernel(KT.PERMUTE_FORWARD, (dims[2], dims[0], dims[1]), 16)
    k1.execute(a, b)
    k2.execute(b, a)
    k3.execute(a, b, sync=True)
    assert same(b.download(inp.size, inp.dtype), inp)


def test_autotuned_kernel_is_exact():
    dims = (96, 80, 16)
    inp = payload(int(np.prod(dims)), 8, seed=4)
    want = np.zeros_like(inp)
    O.host_kernel(KT.PERMUTE_FORWARD, dims, inp, want)
    di = D.DeviceArray.from_numpy(inp)
    do = D.DeviceArray.from_numpy(np.zeros_like(inp))
    k = D.Kernel(KT.PERMUTE_FORWARD, dims, 8, effort=D.Effort.MEASURE)
    k.execute(di, do, sync=True)
    assert same(do.download(inp.size, inp.dtype), want)


def run_offset(kt, dims, es, nd, neighbor, in_shift, out_shift, out_size):
    """run_both with the device pointers moved by whole elements (in_shift / out_shift), so a
    16-byte-lane kernel sees 4- or 8-byte-aligned pointers and has to take its fallback."""
    n = int(np.prod(dims))
    inp = payload(n, es, seed=3)
    want = np.zeros(out_size, dtype=inp.dtype)
    O.host_kernel(kt, dims, inp, want, nd, neighbor or None)
    di = D.DeviceArray.from_numpy(np.concatenate([np.zeros(in_shift, dtype=inp.dtype), inp]))
    do = D.DeviceArray.from_numpy(np.zeros(out_size + out_shift, dtype=inp.dtype))
    k = D.Kernel(kt, dims, es, nd)
    k.execute(di.ptr + in_shift * es, do.ptr + out_shift * es, neighbor=neighbor, sync=True)
    got = do.download(out_size + out_shift, inp.dtype)[out_shift:]
    return got, want


@pytest.mark.parametrize("es", [4, 8])
@pytest.mark.parametrize("case", [
    # (kernel, dims, neighbour rows, neighbour): 16-byte lanes on aligned boxes whose extents
    # end in part of a vector (input side: nxx = 30; output side: UNPACK_FORWARD with nxx = 30),
    # and boxes whose offsets or strides are not whole vectors (per-element fallback)
    (KT.PACK_FORWARD, (96, 40, 12), [[30, 40, 12, 0, 0]], 1),
    (KT.PACK_FORWARD, (96, 40, 12), [[32, 40, 12, 32, 0], [64, 40, 12, 0, 15360]], 2),
    (KT.PACK_FORWARD, (96, 40, 12), [[30, 40, 12, 30, 0]], 1),
    (KT.UNPACK_FORWARD, (96, 40, 12), [[30, 40, 12, 0, 0]], 1),
    (KT.UNPACK_FORWARD, (96, 40, 12), [[36, 40, 12, 0, 60]], 1),
    (KT.PERMUTE_FORWARD, (68, 36, 9), None, 0),
    (KT.PERMUTE_BACKWARD, (36, 9, 68), None, 0),
])
def test_vector_lane_edges_and_fallback(case, es):
    kt, dims, nd, neighbor = case
    n = int(np.prod(dims))
    for shift in (0, 1):   # shift 1: pointers off the 16-byte grid for every element size here
        got, want = run_offset(kt, dims, es, nd, neighbor, shift, shift, n)
        assert same(got, want), (case, es, shift)
