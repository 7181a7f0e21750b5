"""Pins the oracle's transpose / pack / unpack index maps.

The reference stores no golden vectors for these kernels; its unit test
(src/tests/test_host_kernels.F90) checks identities on unique payloads
in(i) = i.  This file runs those identities on the oracle and, independently,
checks every kernel type against numpy axis permutations and slices of the
Fortran-ordered array (a textbook formulation that shares no code with the
oracle's loops).  Shapes are the reference's: [33,77,21], [90,57] (host test)
and [18,33,155], [18,155] (device test).
"""
import numpy as np
import pytest

import oracle as O

SHAPES3 = [(33, 77, 21), (18, 33, 155), (5, 1, 7)]
SHAPES2 = [(90, 57), (18, 155), (1, 9)]
ES = [4, 8, 16]


def payload(n, es):
    """Unique payloads (the reference uses in(i) = i), stored as raw es-byte items."""
    a = np.arange(1, n + 1, dtype=np.uint64)
    if es == 4:
        return a.astype(np.uint32)
    if es == 8:
        return a
    v = np.empty(n, dtype=O.elem_dtype(16))
    raw = v.view(np.uint64).reshape(n, 2)
    raw[:, 0] = a
    raw[:, 1] = a * 1000003
    return v


def run(kt, dims, inp, out_size=None, nd=None, neighbor=None, access="write"):
    out = np.zeros(out_size or inp.size, dtype=inp.dtype)
    O.host_kernel(kt, dims, inp, out, nd, neighbor, access)
    return out


def f3(a, dims):
    """View a flat buffer as a Fortran [d0, d1, d2] array (numpy shape (d2, d1, d0))."""
    return a.reshape(dims[2], dims[1], dims[0])


@pytest.mark.parametrize("es", ES)
@pytest.mark.parametrize("dims", SHAPES3)
def test_permutes_3d_against_numpy(es, dims):
    nx, ny, nz = dims
    inp = payload(nx * ny * nz, es)
    a = f3(inp, dims)  # a[z, y, x]
    fwd = run(O.KERNEL_PERMUTE_FORWARD, dims, inp)
    assert np.array_equal(fwd, np.transpose(a, (2, 0, 1)).ravel())   # [Y, Z, X]
    bwd = run(O.KERNEL_PERMUTE_BACKWARD, dims, inp)
    assert np.array_equal(bwd, np.transpose(a, (1, 2, 0)).ravel())   # [Z, X, Y]
    bst = run(O.KERNEL_PERMUTE_BACKWARD_START, dims, inp)
    assert np.array_equal(bst, np.transpose(a, (2, 1, 0)).ravel())   # [Z, Y, X]
    for kt in (O.KERNEL_PERMUTE_FORWARD, O.KERNEL_PERMUTE_BACKWARD, O.KERNEL_PERMUTE_BACKWARD_START):
        assert np.array_equal(run(kt, dims, inp, access="read"), run(kt, dims, inp)), kt


@pytest.mark.parametrize("es", ES)
@pytest.mark.parametrize("dims", SHAPES2)
def test_permutes_2d_against_numpy(es, dims):
    nx, ny = dims
    inp = payload(nx * ny, es)
    a = inp.reshape(ny, nx)
    want = a.T.ravel()
    assert np.array_equal(run(O.KERNEL_PERMUTE_FORWARD, dims, inp), want)
    # 2D BACKWARD is remapped to FORWARD (src/dtfft_abstract_kernel.F90:272-283)
    assert np.array_equal(run(O.KERNEL_PERMUTE_BACKWARD, dims, inp), want)
    assert np.array_equal(run(O.KERNEL_PERMUTE_FORWARD, dims, inp, access="read"), want)


@pytest.mark.parametrize("dims", SHAPES3)
def test_identities_3d(dims):
    """forward^3 = id, backward o forward = id, backward_end o backward_start = backward."""
    nx, ny, nz = dims
    inp = payload(nx * ny * nz, 8)
    a1 = run(O.KERNEL_PERMUTE_FORWARD, (nx, ny, nz), inp)
    a2 = run(O.KERNEL_PERMUTE_FORWARD, (ny, nz, nx), a1)
    a3 = run(O.KERNEL_PERMUTE_FORWARD, (nz, nx, ny), a2)
    assert np.array_equal(a3, inp)
    b1 = run(O.KERNEL_PERMUTE_BACKWARD, (ny, nz, nx), a1)
    assert np.array_equal(b1, inp)
    # backward_end(backward_start(x)) == backward(x), src/tests/test_host_kernels.F90:195-234
    gold = run(O.KERNEL_PERMUTE_BACKWARD, dims, inp)
    tmp = run(O.KERNEL_PERMUTE_BACKWARD_START, dims, inp)
    tdims = (nz, nx, ny)
    nd = np.array([[nz, nx, ny, 0, 0]])
    for access in ("write", "read"):
        out = run(O.KERNEL_PERMUTE_BACKWARD_END, tdims, tmp, nd=nd, access=access)
        assert np.array_equal(out, gold)
        out = run(O.KERNEL_PERMUTE_BACKWARD_END_PIPELINED, tdims, tmp, nd=nd, neighbor=1, access=access)
        assert np.array_equal(out, gold)


@pytest.mark.parametrize("dims", SHAPES3 + SHAPES2)
def test_pack_unpack_identity(dims):
    """pack/unpack with din = dout = 0 and full dims is the identity (test_host_kernels.F90:236-272)."""
    inp = payload(int(np.prod(dims)), 4)
    nd = np.array([list(dims) + [1] * (3 - len(dims)) + [0, 0]])
    for kt in (O.KERNEL_PACK, O.KERNEL_UNPACK):
        assert np.array_equal(run(kt, dims, inp, nd=nd), inp)
    for kt in (O.KERNEL_PACK_PIPELINED, O.KERNEL_UNPACK_PIPELINED):
        assert np.array_equal(run(kt, dims, inp, nd=nd, neighbor=1), inp)


@pytest.mark.parametrize("dims", SHAPES3 + SHAPES2)
def test_fused_pack_and_unpack_forward_equal_permute(dims):
    """pack_forward with full dims == permute_forward; unpack_forward with the transposed
    dims == permute_forward (test_host_kernels.F90:48-95)."""
    inp = payload(int(np.prod(dims)), 4)
    gold = run(O.KERNEL_PERMUTE_FORWARD, dims, inp)
    nd = np.array([list(dims) + [1] * (3 - len(dims)) + [0, 0]])
    for access in ("write", "read"):
        assert np.array_equal(run(O.KERNEL_PACK_FORWARD, dims, inp, nd=nd, neighbor=1, access=access), gold)
    udims = (dims[1], dims[0]) if len(dims) == 2 else (dims[1], dims[2], dims[0])
    und = np.array([list(udims) + [1] * (3 - len(dims)) + [0, 0]])
    for access in ("write", "read"):
        assert np.array_equal(run(O.KERNEL_UNPACK_FORWARD, udims, inp, nd=und, access=access), gold)


@pytest.mark.parametrize("dims", SHAPES3)
def test_fused_pack_and_unpack_backward_equal_permute(dims):
    inp = payload(int(np.prod(dims)), 4)
    gold = run(O.KERNEL_PERMUTE_BACKWARD, dims, inp)
    nd = np.array([list(dims) + [0, 0]])
    for access in ("write", "read"):
        assert np.array_equal(run(O.KERNEL_PACK_BACKWARD, dims, inp, nd=nd, neighbor=1, access=access), gold)
    udims = (dims[2], dims[0], dims[1])
    und = np.array([list(udims) + [0, 0]])
    for access in ("write", "read"):
        assert np.array_equal(run(O.KERNEL_UNPACK_BACKWARD, udims, inp, nd=und, access=access), gold)


def test_pack_forward_subbox_against_numpy():
    """Per-peer pack_forward on an x-offset sub-box equals slice + axis permutation."""
    nx, ny, nz = 40, 6, 5
    inp = payload(nx * ny * nz, 8)
    a = f3(inp, (nx, ny, nz))
    out_size = nx * ny * nz
    nd = []
    dout = 0
    for x0, w in ((0, 10), (10, 10), (20, 20)):
        nd.append([w, ny, nz, x0, dout])
        dout += w * ny * nz
    nd = np.array(nd)
    out = np.zeros(out_size, dtype=inp.dtype)
    for peer in range(1, 4):
        O.host_kernel(O.KERNEL_PACK_FORWARD, (nx, ny, nz), inp, out, nd, peer)
    want = np.concatenate([np.transpose(a[:, :, x0:x0 + w], (2, 0, 1)).ravel()
                           for x0, w in ((0, 10), (10, 10), (20, 20))])
    assert np.array_equal(out, want)


def test_pack_unpack_subbox_against_numpy():
    """pack_pipelined gathers x-slabs into contiguous chunks; unpack scatters them back."""
    nx, ny, nz = 24, 7, 3
    inp = payload(nx * ny * nz, 4)
    a = f3(inp, (nx, ny, nz))
    parts = ((0, 8), (8, 5), (13, 11))
    nd, d = [], 0
    for x0, w in parts:
        nd.append([w, ny, nz, x0, d])
        d += w * ny * nz
    nd = np.array(nd)
    packed = run(O.KERNEL_PACK, (nx, ny, nz), inp, nd=nd)
    want = np.concatenate([a[:, :, x0:x0 + w].ravel() for x0, w in parts])
    assert np.array_equal(packed, want)
    # unpack: in offsets are the chunk offsets, out offsets the x starts
    und = np.array([[w, ny, nz, row[4], x0] for (x0, w), row in zip(parts, nd)])
    back = run(O.KERNEL_UNPACK, (nx, ny, nz), packed, nd=und)
    assert np.array_equal(back, inp)


def test_copy_kernels():
    inp = payload(100, 16)
    assert np.array_equal(run(O.KERNEL_COPY, (10, 10), inp), inp)
    nd = np.array([[5, 2, 1, 10, 30]])
    out = run(O.KERNEL_COPY_PIPELINED, (10, 10), inp, nd=nd, neighbor=1)
    assert np.array_equal(out[30:40], inp[10:20])
    assert (out.view(np.uint64)[:60] == 0).all()


def test_empty_dims_is_dummy():
    inp = payload(10, 4)
    out = run(O.KERNEL_PERMUTE_FORWARD, (10, 0, 1), inp)
    assert (out == 0).all()


@pytest.mark.parametrize("es", ES)
@pytest.mark.parametrize("dims", SHAPES3 + SHAPES2)
def test_blocked_host_variants_against_numpy(es, dims):
    """The cache-blocked BLOCK_4..64 host permutes (_dtfft_kernel_host_block_routines.inc),
    in both loop orders, give the numpy axis permutation for every block size, including
    blocks that do not divide the extents."""
    inp = payload(int(np.prod(dims)), es)
    if len(dims) == 3:
        a = f3(inp, dims)
        want = {O.KERNEL_PERMUTE_FORWARD: np.transpose(a, (2, 0, 1)).ravel(),
                O.KERNEL_PERMUTE_BACKWARD: np.transpose(a, (1, 2, 0)).ravel()}
    else:
        a = inp.reshape(dims[1], dims[0])
        want = {O.KERNEL_PERMUTE_FORWARD: a.T.ravel()}
    for kt, w in want.items():
        for block in O.BLOCK_SIZES:
            for access in ("write", "read"):
                out = np.zeros_like(inp)
                O.permute_variant(kt, dims, inp, out, block, access)
                assert np.array_equal(out, w), (kt, block, access)


def test_host_autotune_picks_a_listed_variant_that_is_exact():
    """The timed selection (kernel_host%create_host at EXHAUSTIVE) returns BASE or a BLOCK
    size with an access mode, and the plan-level cycle run with the picked variants equals
    the BASE run bit for bit."""
    import oracle.plan as OP
    dims = (33, 20, 14)
    tuned = OP.autotune_generic(dims, 8, warmup=1, iters=1)
    assert len(tuned) == 4
    for block, access in tuned.values():
        assert block in (0,) + O.BLOCK_SIZES and access in ("write", "read")
    rng = np.random.default_rng(0)
    x = (rng.uniform(-1, 1, dims[::-1]) + 1j * rng.uniform(-1, 1, dims[::-1])).astype(np.complex64)
    z0 = OP.execute_generic(x, True)
    z1 = OP.execute_generic(x, True, kernels=tuned)
    assert z0.tobytes() == z1.tobytes()
    assert OP.execute_generic(z0, False, kernels=tuned).tobytes() == OP.execute_generic(z0, False).tobytes()
