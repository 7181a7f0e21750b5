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
