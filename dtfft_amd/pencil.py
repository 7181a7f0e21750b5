"""Pencil geometry and per-peer neighbour data of a P > 1 transpose.

On several ranks a forward transpose (X->Y, Y->Z) exchanges, inside a 1D
sub-communicator of P peers, the blocks of a pencil A = [a, b_me, c] (a complete,
b split across the peers) to build pencil B = [b, c, a_me] (b complete, a split).
The kernel plugin receives, per peer, a row (nxx, nyy, nzz, din, dout):

* send side (src/dtfft_reshape_handle_generic.F90:291-423, transposes): peer i gets
  the a-range of its B pencil: (a_i, b_me, c, a_start_i, running send displacement);
  the fused backends pack it with KERNEL_PACK_FORWARD (get_fused,
  src/dtfft_abstract_kernel.F90:202-217), so the chunk is laid out [b_me, c, a_i];
* receive side (:485-616): the chunk from peer i holds (b_i, c, a_me) and lands at the
  b-offset of peer i: (b_i, c, a_me, running receive displacement, b_start_i), run by
  KERNEL_UNPACK(_PIPELINED).

Counts split a global extent as src/dtfft_pencil.F90:237-279 does (the remainder goes
to the last ranks).  This module is host logic only; the data movement is done by the
HIP kernels.
"""
from __future__ import annotations


def local_size(n_global: int, comm_size: int, rank: int) -> tuple[int, int]:
    """(start, count) of `rank` in a 1D split of n_global (src/dtfft_pencil.F90:237-279)."""
    if comm_size == 1:
        return 0, n_global
    res = n_global % comm_size
    counts = [n_global // comm_size + (1 if r >= comm_size - res else 0) for r in range(comm_size)]
    return sum(counts[:rank]), counts[rank]


def split(n_global: int, comm_size: int) -> list[tuple[int, int]]:
    return [local_size(n_global, comm_size, r) for r in range(comm_size)]


def forward_exchange(a_global: int, b_global: int, c_local: int, comm_size: int, me: int):
    """Neighbour data rows of a forward transpose A=[a, b_me, c] -> B=[b, c, a_me].

    Returns (send_rows, recv_rows, send_counts, recv_counts); rows are
    (nxx, nyy, nzz, din, dout) per peer in peer order, counts in elements."""
    a_parts = split(a_global, comm_size)   # a-ranges of the peers' B pencils
    b_parts = split(b_global, comm_size)   # b-ranges of the peers' A pencils
    b_me = b_parts[me][1]
    a_me = a_parts[me][1]
    send, recv, scount, rcount = [], [], [], []
    sdispl = rdispl = 0
    for i in range(comm_size):
        a_start, a_cnt = a_parts[i]
        n = a_cnt * b_me * c_local
        send.append((a_cnt, b_me, c_local, a_start, sdispl))
        scount.append(n)
        sdispl += n
        b_start, b_cnt = b_parts[i]
        m = b_cnt * c_local * a_me
        recv.append((b_cnt, c_local, a_me, rdispl, b_start))
        rcount.append(m)
        rdispl += m
    return send, recv, scount, rcount


# ---------------------------------------------------------------------------------------
# Every transpose of a P > 1 plan, both directions, uneven splits
# ---------------------------------------------------------------------------------------
# dtfft_transpose_t values (include/dtfft_config.h.in:35-43)
X_TO_Y, Y_TO_X, Y_TO_Z, Z_TO_Y, X_TO_Z, Z_TO_X = 1, -1, 2, -2, 3, -3
FORWARD_TYPES = (X_TO_Y, Y_TO_Z, Z_TO_X)

# kernel_type_t values (src/dtfft_abstract_kernel.F90:59-98)
KERNEL_PACK, KERNEL_UNPACK = 1, 3
KERNEL_UNPACK_PIPELINED, KERNEL_PACK_PIPELINED = 5, 6
KERNEL_PERMUTE_FORWARD, KERNEL_PERMUTE_BACKWARD = 7, 8
KERNEL_PERMUTE_BACKWARD_START, KERNEL_PERMUTE_BACKWARD_END = 9, 10
KERNEL_PERMUTE_BACKWARD_END_PIPELINED = 11
KERNEL_PACK_FORWARD, KERNEL_PACK_BACKWARD = 12, 13
KERNEL_UNPACK_FORWARD, KERNEL_UNPACK_FORWARD_PIPELINED = 15, 16
KERNEL_UNPACK_BACKWARD, KERNEL_UNPACK_BACKWARD_PIPELINED = 17, 18

# How the exchange is driven (the backend families of src/dtfft_reshape_handle_generic.F90:
# 401-413, 582-600): "mpi" packs the whole block, exchanges, then unpacks every peer;
# "pipelined" unpacks peer by peer as chunks arrive; "fused" packs peer by peer straight
# into the send chunks (get_fused) and unpacks peer by peer; "unpack" is the host's
# transpose mode UNPACK (plain per-peer copies out, the permutation on the receive side).
MODES = ("mpi", "pipelined", "fused", "unpack")


def sub_pencils(transpose_type: int, dims, comm_size: int, fixed=(None, 0)):
    """Send and receive pencils of every rank of the 1D sub-communicator a transpose runs in.

    ``dims`` are the global dims (X first).  Pencils follow get_permutations and
    create_pencils_and_comm (src/dtfft_transpose_plan.F90:1045-1131): X = [X, Y, Z],
    Y = [Y, Z, X], Z = [Z, X, Y] (2D: X = [X, Y], Y = [Y, X]).  X<->Y exchanges inside the
    communicator that splits Y of the X-pencil, Y<->Z inside the one that splits Z, X<->Z
    (Z-slab) inside the one that splits Z while Y is whole.  The axis that the transpose
    leaves alone keeps this rank's (count, start) = ``fixed`` (count None: the whole axis).
    Returns a list over ranks of (in_counts, in_starts, out_counts, out_starts), each in the
    pencil's own axis order."""
    nd = len(dims)
    up = transpose_type > 0  # X->Y, Y->Z, X->Z run from the lower pencil to the higher one
    tt = abs(transpose_type)
    res = []
    for r in range(comm_size):
        if nd == 2:
            nx, ny = dims
            ys, yn = local_size(ny, comm_size, r)
            xs, xn = local_size(nx, comm_size, r)
            xp = ((nx, yn), (0, ys))
            yp = ((ny, xn), (0, xs))
            a, b = (xp, yp) if tt == X_TO_Y else (None, None)
        else:
            nx, ny, nz = dims
            if tt == X_TO_Y:
                fc = nz if fixed[0] is None else fixed[0]
                ys, yn = local_size(ny, comm_size, r)
                xs, xn = local_size(nx, comm_size, r)
                a = ((nx, yn, fc), (0, ys, fixed[1]))
                b = ((ny, fc, xn), (0, fixed[1], xs))
            elif tt == Y_TO_Z:
                fc = nx if fixed[0] is None else fixed[0]
                zs, zn = local_size(nz, comm_size, r)
                ys, yn = local_size(ny, comm_size, r)
                a = ((ny, zn, fc), (0, zs, fixed[1]))
                b = ((nz, fc, yn), (0, fixed[1], ys))
            else:  # X_TO_Z, Z-slab: Y whole in the X-pencil, X whole in the Z-pencil
                zs, zn = local_size(nz, comm_size, r)
                ys, yn = local_size(ny, comm_size, r)
                a = ((nx, ny, zn), (0, 0, zs))
                b = ((nz, nx, yn), (0, 0, ys))
        if nd == 2 and tt != X_TO_Y:
            raise ValueError("2D plans have X<->Y transposes only")
        send, recv = (a, b) if up else (b, a)
        res.append((send[0], send[1], recv[0], recv[1]))
    return res


def transpose_neighbor_data(transpose_type: int, ndims: int, pencils, me: int):
    """neighbor_data of the pack and unpack kernels of one rank: the transpose branch of
    reshape_handle_generic%create, send side src/dtfft_reshape_handle_generic.F90:291-423 and
    receive side :485-616.  ``pencils`` is sub_pencils' list (every rank's send and receive
    pencil, as the handles' MPI_Allgather collects them, :136-145).

    Returns (send_rows, recv_rows, send_counts, recv_counts): rows (nxx, nyy, nzz, din, dout)
    per peer in peer order, counts in elements (the MPI_Alltoallv counts)."""
    P = len(pencils)
    in_sizes = [p[0] for p in pencils]
    in_starts = [p[1] for p in pencils]
    out_sizes = [p[2] for p in pencils]
    out_starts = [p[3] for p in pencils]

    def send_ln_ls(src: int, i: int):
        """in%ln(:, i), in%ls(:, i) as computed on rank `src` for peer i (:295-334)."""
        if ndims == 2:
            return (out_sizes[i][1], in_sizes[src][1]), (out_starts[i][1], in_starts[src][1])
        if transpose_type == X_TO_Z:
            return ((in_sizes[src][0], out_sizes[i][2], in_sizes[src][2]),
                    (in_starts[src][0], out_starts[i][2], in_starts[src][2]))
        if transpose_type in (Z_TO_X, X_TO_Y, Y_TO_Z):
            return ((out_sizes[i][2], in_sizes[src][1], in_sizes[src][2]),
                    (out_starts[i][2], in_starts[src][1], in_starts[src][2]))
        return ((out_sizes[i][1], in_sizes[src][1], in_sizes[src][2]),
                (out_starts[i][1], in_starts[src][1], in_starts[src][2]))

    def prod(v):
        n = 1
        for x in v:
            n *= int(x)
        return n

    send, scount = [], []
    sdispl = 0
    for i in range(P):
        ln, ls = send_ln_ls(me, i)
        din = ln[0] * ls[1] if transpose_type == X_TO_Z else ls[0]
        send.append((ln[0], ln[1], ln[2] if ndims == 3 else 1, din, sdispl))
        n = prod(ln)
        scount.append(n)
        sdispl += n

    recv, rcount = [], []
    rdispl = 0
    for i in range(P):
        recvsize = prod(send_ln_ls(i, me)[0])  # what rank i sends here (:487-488)
        ln, ls = (0, 0, 0), (0, 0, 0)
        if recvsize > 0:
            if ndims == 2:
                ln = (in_sizes[i][1], out_sizes[me][1], 1)
                ls = (in_starts[i][1], out_starts[me][1], 0)
            elif transpose_type == X_TO_Z:
                ln = (in_sizes[i][2], out_sizes[me][1], out_sizes[me][2])
                ls = (in_starts[i][2], out_starts[me][1], out_starts[me][2])
            elif transpose_type == Z_TO_X:
                ln = (out_sizes[me][0], in_sizes[i][2], out_sizes[me][2])
                ls = (out_starts[me][0], in_starts[i][2], out_starts[me][2])
            elif transpose_type in (X_TO_Y, Y_TO_Z):
                ln = (in_sizes[i][1], out_sizes[me][1], out_sizes[me][2])
                ls = (in_starts[i][1], in_starts[me][1], out_starts[me][2])
            else:
                ln = (in_sizes[i][2], out_sizes[me][1], out_sizes[me][2])
                ls = (in_starts[i][2], in_starts[me][1], out_starts[me][2])
        dout = ln[0] * ls[1] if transpose_type == Z_TO_X else ls[0]
        recv.append((ln[0], ln[1], ln[2] if ndims == 3 else 1, rdispl, dout))
        rcount.append(recvsize)
        rdispl += recvsize
    return send, recv, scount, rcount


def transpose_kernel_types(transpose_type: int, ndims: int, mode: str):
    """(pack kernel, unpack kernel, two_step) of a P > 1 transpose, as
    src/dtfft_reshape_handle_generic.F90:236-245 (permute direction), :401-413 (two-step
    backward, transpose mode UNPACK, get_fused) and :582-600 (unpack side) select them.
    dtFFT on CUDA ignores the UNPACK transpose mode (:211); here it is kept for the host
    restatement and the GPU kernels alike."""
    if mode not in MODES:
        raise ValueError(f"mode must be one of {MODES}")
    pack = KERNEL_PERMUTE_FORWARD if transpose_type in FORWARD_TYPES else KERNEL_PERMUTE_BACKWARD
    fused = mode == "fused"
    pipelined = mode in ("pipelined", "fused")
    two_step = transpose_type in (Y_TO_X, Z_TO_Y) and ndims == 3 and not fused
    if two_step:
        pack = KERNEL_PERMUTE_BACKWARD_START
    if mode == "unpack":
        pack = KERNEL_PACK
    if fused:
        pack = {KERNEL_PACK: KERNEL_PACK_PIPELINED, KERNEL_PERMUTE_FORWARD: KERNEL_PACK_FORWARD,
                KERNEL_PERMUTE_BACKWARD: KERNEL_PACK_BACKWARD}.get(pack, pack)
    unpack = KERNEL_UNPACK_PIPELINED if pipelined else KERNEL_UNPACK
    if two_step:
        unpack = KERNEL_PERMUTE_BACKWARD_END_PIPELINED if pipelined else KERNEL_PERMUTE_BACKWARD_END
    if mode == "unpack":
        if transpose_type in FORWARD_TYPES:
            unpack = KERNEL_UNPACK_FORWARD_PIPELINED if pipelined else KERNEL_UNPACK_FORWARD
        else:
            unpack = KERNEL_UNPACK_BACKWARD_PIPELINED if pipelined else KERNEL_UNPACK_BACKWARD
    return pack, unpack, two_step


class TransposeShare:
    """One rank's P > 1 transpose: pencils, neighbour data, kernel types and the
    pack -> exchange -> unpack order of reshape_handle_generic%execute
    (src/dtfft_reshape_handle_generic.F90:695-759).  Host logic only: ``run`` drives the
    caller's kernels (host or GPU) through three callables."""

    def __init__(self, transpose_type: int, dims, comm_size: int, me: int, mode: str = "mpi", fixed=(None, 0)):
        self.transpose_type, self.dims, self.P, self.me, self.mode = transpose_type, tuple(dims), comm_size, me, mode
        self.ndims = len(dims)
        self.pencils = sub_pencils(transpose_type, dims, comm_size, fixed)
        self.in_counts, self.in_starts, self.out_counts, self.out_starts = self.pencils[me]
        (self.send_rows, self.recv_rows, self.send_counts,
         self.recv_counts) = transpose_neighbor_data(transpose_type, self.ndims, self.pencils, me)
        self.pack_kernel, self.unpack_kernel, self.two_step = transpose_kernel_types(transpose_type, self.ndims,
                                                                                     mode)
        self.send_displs = [r[4] for r in self.send_rows]
        self.recv_displs = [r[3] for r in self.recv_rows]

    @property
    def per_peer_pack(self) -> bool:
        return self.pack_kernel in (KERNEL_PACK_FORWARD, KERNEL_PACK_BACKWARD, KERNEL_PACK_PIPELINED)

    @property
    def per_peer_unpack(self) -> bool:
        return self.mode in ("pipelined", "fused")

    def run_pack(self, pack):
        """pack(neighbor) for each peer (1-based) when the pack kernel is per peer, else pack(0)
        once (a whole-block permute, or PACK over every peer)."""
        if self.per_peer_pack:
            for i in range(1, self.P + 1):
                pack(i)
        else:
            pack(0)

    def run_unpack(self, unpack):
        if self.per_peer_unpack:
            for i in range(1, self.P + 1):
                unpack(i)
        else:
            unpack(0)


# ---------------------------------------------------------------------------------------
# Bricks <-> pencils reshapes (src/dtfft_reshape_plan.F90:98-260, the reshape branches of
# src/dtfft_reshape_handle_generic.F90:261-289, 343-403, 535-565, 584-596)
# ---------------------------------------------------------------------------------------
# dtfft_reshape_t values (include/dtfft_config.h.in:45-48)
X_BRICKS_TO_PENCILS, X_PENCILS_TO_BRICKS, Z_PENCILS_TO_BRICKS, Z_BRICKS_TO_PENCILS = 11, 12, 13, 14
BRICKS_TO_PENCILS = (X_BRICKS_TO_PENCILS, Z_BRICKS_TO_PENCILS)
KERNEL_DUMMY = -1
RESHAPE_MODES = ("mpi", "pipelined", "fused")


def sub_bricks(reshape_type: int, dims, comm_size: int, split=None):
    """Bricks and pencils of every rank of the sub-communicator a reshape runs in.

    ``dims`` are the group's extents in the pencil's own axis order (X-pencil [X, Y, Z] for the
    X reshapes, Z-pencil [Z, X, Y] for the Z reshapes; 2D [X, Y]).  In the bricks the first
    axis is split over the ``comm_size`` ranks and the others are whole; in the pencils the
    first axis is whole and the others are split ``split = (a, b)`` ways (a * b = comm_size;
    2D: a = comm_size), rank i holding part (i % a, i // a): the order of the group's ranks by
    pencil start that the reshape plan's MPI_Comm_split gives them
    (src/dtfft_reshape_plan.F90:168-174).  Returns the list over ranks of (in_counts,
    in_starts, out_counts, out_starts) in that axis order."""
    nd = len(dims)
    P = comm_size
    if split is None:
        split = (P, 1) if nd == 2 else (1, P)
    a, b = split
    if a * b != P or (nd == 2 and b != 1):
        raise ValueError("split must multiply to the group size (2D: (P, 1))")
    res = []
    for r in range(P):
        xs, xn = local_size(dims[0], P, r)
        ys, yn = local_size(dims[1], a, r % a)
        if nd == 2:
            brick = ((xn, dims[1]), (xs, 0))
            pen = ((dims[0], yn), (0, ys))
        else:
            zs, zn = local_size(dims[2], b, r // a)
            brick = ((xn, dims[1], dims[2]), (xs, 0, 0))
            pen = ((dims[0], yn, zn), (0, ys, zs))
        send, recv = (brick, pen) if reshape_type in BRICKS_TO_PENCILS else (pen, brick)
        res.append((send[0], send[1], recv[0], recv[1]))
    return res


def reshape_neighbor_data(reshape_type: int, ndims: int, pencils, me: int):
    """neighbor_data of a reshape's pack and unpack kernels on one rank, the reshape branches
    of reshape_handle_generic%create: send side src/dtfft_reshape_handle_generic.F90:343-403,
    receive side :535-565 and :584-596, the strategy of :265-289 (1: every pencil holds the
    bricks' whole second axis, 2: the whole third axis, 3: neither) and the pack-free /
    unpack-free tests of :261-264 and :479-484.

    Returns (send_rows, recv_rows, send_counts, recv_counts, strat, pack_free, unpack_free)."""
    P = len(pencils)
    in_sizes = [p[0] for p in pencils]
    in_starts = [p[1] for p in pencils]
    out_sizes = [p[2] for p in pencils]
    out_starts = [p[3] for p in pencils]
    b2p = reshape_type in BRICKS_TO_PENCILS
    # pack-free (bricks -> pencils) / unpack-free (pencils -> bricks): every peer's pencil part
    # holds the whole second axis of my brick (the test of every rank agrees, :259-264)
    second_whole = all(in_sizes[me][1] == out_sizes[i][1] for i in range(P)) or ndims == 2
    pack_free = b2p and second_whole
    unpack_free = (not b2p) and second_whole
    if ndims == 2:
        strat = 1
    else:
        if b2p:
            zslab = all(in_sizes[me][1] == out_sizes[i][1] for i in range(P))
            yslab = all(in_sizes[me][2] == out_sizes[i][2] for i in range(P))
        else:
            zslab = all(in_sizes[i][1] == out_sizes[me][1] for i in range(P))
            yslab = all(in_sizes[i][2] == out_sizes[me][2] for i in range(P))
        strat = 1 if zslab else (2 if yslab else 3)

    def send_ln_ls(src: int, i: int):
        if ndims == 2:
            if b2p:
                return (in_sizes[src][0], out_sizes[i][1]), (in_starts[src][0], out_starts[i][1])
            return (out_sizes[i][0], in_sizes[src][1]), (out_starts[i][0], in_starts[src][1])
        if b2p:
            return ((in_sizes[src][0], out_sizes[i][1], out_sizes[i][2]),
                    (in_starts[src][0], out_starts[i][1], out_starts[i][2]))
        return ((out_sizes[i][0], in_sizes[src][1], in_sizes[src][2]),
                (out_starts[i][0], in_starts[src][1], in_starts[src][2]))

    def prod(v):
        n = 1
        for x in v:
            n *= int(x)
        return n

    send, scount = [], []
    sdispl = ssdispl = 0
    sc, ss = in_sizes[me], in_starts[me]
    for i in range(P):
        ln, ls = send_ln_ls(me, i)
        if b2p:
            if strat == 1:
                din = ssdispl
                ssdispl += prod(ln)
            elif strat == 2:
                din = ssdispl
                ssdispl += ln[0] * ln[1]
            else:
                din = (out_starts[i][2] - ss[2]) * sc[0] * sc[1] + (out_starts[i][1] - ss[1]) * sc[0]
        else:
            din = ls[0]
        send.append((ln[0], ln[1], ln[2] if ndims == 3 else 1, din, sdispl))
        n = prod(ln)
        scount.append(n)
        sdispl += n

    recv, rcount = [], []
    rdispl = rrdispl = 0
    rc, rs = out_sizes[me], out_starts[me]
    for i in range(P):
        recvsize = prod(send_ln_ls(i, me)[0])
        ln = (0, 0, 0)
        ls = (0, 0, 0)
        if recvsize > 0:
            if ndims == 2:
                if b2p:
                    ln, ls = (in_sizes[i][0], out_sizes[me][1], 1), (in_starts[i][0], out_starts[me][1], 0)
                else:
                    ln, ls = (out_sizes[me][0], in_sizes[i][1], 1), (out_starts[me][0], in_starts[i][1], 0)
            elif b2p:
                ln = (in_sizes[i][0], out_sizes[me][1], out_sizes[me][2])
                ls = (in_starts[i][0], out_starts[me][1], out_starts[me][2])
            else:
                ln = (out_sizes[me][0], in_sizes[i][1], in_sizes[i][2])
                ls = (out_starts[me][0], in_starts[i][1], in_starts[i][2])
        if b2p:
            dout = ls[0]
        elif strat in (1, 2):
            dout = rrdispl
            rrdispl += ln[0] * ln[1]
        else:
            dout = abs(rs[2] - in_starts[i][2]) * rc[0] * rc[1] + abs(rs[1] - in_starts[i][1]) * rc[0]
        recv.append((ln[0], ln[1], ln[2] if ndims == 3 else 1, rdispl, dout))
        rcount.append(recvsize)
        rdispl += recvsize
    return send, recv, scount, rcount, strat, pack_free, unpack_free


def reshape_kernel_types(mode: str, pack_free: bool, unpack_free: bool):
    """(pack kernel, unpack kernel) of a reshape (src/dtfft_reshape_handle_generic.F90:415-419,
    586-595): PACK / UNPACK, their per-peer forms when fused / pipelined, DUMMY where the
    layout makes them unnecessary."""
    if mode not in RESHAPE_MODES:
        raise ValueError(f"mode must be one of {RESHAPE_MODES}")
    pack = KERNEL_DUMMY if pack_free else (KERNEL_PACK_PIPELINED if mode == "fused" else KERNEL_PACK)
    unpack = KERNEL_DUMMY if unpack_free else (KERNEL_UNPACK_PIPELINED if mode in ("pipelined", "fused")
                                               else KERNEL_UNPACK)
    return pack, unpack


class ReshapeShare:
    """One rank's bricks <-> pencils reshape: geometry, neighbour data, kernel types, and the
    pack -> exchange -> unpack order of reshape_handle_generic%execute
    (src/dtfft_reshape_handle_generic.F90:695-759).  With a DUMMY pack (pack-free) the input
    is the send buffer; with a DUMMY unpack (unpack-free) the receive buffer is the output."""

    def __init__(self, reshape_type: int, dims, comm_size: int, me: int, mode: str = "mpi", split=None):
        self.reshape_type, self.dims, self.P, self.me, self.mode = reshape_type, tuple(dims), comm_size, me, mode
        self.ndims = len(dims)
        self.pencils = sub_bricks(reshape_type, dims, comm_size, split)
        self.in_counts, self.in_starts, self.out_counts, self.out_starts = self.pencils[me]
        (self.send_rows, self.recv_rows, self.send_counts, self.recv_counts, self.strat, self.pack_free,
         self.unpack_free) = reshape_neighbor_data(reshape_type, self.ndims, self.pencils, me)
        self.pack_kernel, self.unpack_kernel = reshape_kernel_types(mode, self.pack_free, self.unpack_free)
        self.send_displs = [r[4] for r in self.send_rows]
        self.recv_displs = [r[3] for r in self.recv_rows]

    def run_pack(self, pack):
        if self.pack_kernel == KERNEL_DUMMY:
            return
        if self.pack_kernel == KERNEL_PACK_PIPELINED:
            for i in range(1, self.P + 1):
                pack(i)
        else:
            pack(0)

    def run_unpack(self, unpack):
        if self.unpack_kernel == KERNEL_DUMMY:
            return
        if self.unpack_kernel == KERNEL_UNPACK_PIPELINED:
            for i in range(1, self.P + 1):
                unpack(i)
        else:
            unpack(0)
