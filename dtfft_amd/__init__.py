"""dtfft_amd -- MI355X (gfx950) implementation of dtFFT's single-GPU hot path.

Python mirror of the dtFFT plugin and plan interfaces over the C-ABI of
``libdtfft_hip.so`` (``include/dtfft_hip.h``):

* :class:`Kernel`   -- kernel plugin (permute / pack / unpack), dtFFT ``kernel_device``;
* :class:`Executor` -- FFT executor plugin (batched 1D/2D C2C, R2C/C2R), dtFFT
  ``abstract_executor``;
* :class:`PlanC2C`, :class:`PlanR2C` -- single-rank plans with dtFFT's Python API
  names (``execute``, ``transpose``, ``local_sizes``, ``alloc_size``, ...);
* :class:`Cycle`    -- many independent grids in one persistent launch per stage;
* :class:`DeviceArray` -- a device buffer with numpy upload / download.

All compute runs in hand-written HIP kernels compiled by hipRTC for gfx950; there is
no CPU fallback (the library raises if it is not built or no gfx950 GPU is present).
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from ._lib import Config, DtfftError, call, check, lib  # noqa: F401

__all__ = ["Precision", "Effort", "Executor", "Execute", "Transpose", "KernelType", "FFTType",
           "DeviceArray", "Kernel", "Executor", "FFTExecutor", "PlanC2C", "PlanR2C", "Cycle",
           "Config", "DtfftError", "set_config", "device_arch", "synchronize"]


class Precision(enum.IntEnum):
    SINGLE = 0
    DOUBLE = 1


class Effort(enum.IntEnum):
    ESTIMATE = 0
    MEASURE = 1
    PATIENT = 2
    EXHAUSTIVE = 3


class Executor(enum.IntEnum):
    NONE = 0
    HIP = 5


class Execute(enum.IntEnum):
    FORWARD = 11
    BACKWARD = 12


class Transpose(enum.IntEnum):
    X_TO_Y = 1
    Y_TO_X = -1
    Y_TO_Z = 2
    Z_TO_Y = -2
    X_TO_Z = 3
    Z_TO_X = -3


class KernelType(enum.IntEnum):
    """kernel_type_t, src/dtfft_abstract_kernel.F90:59-98"""
    DUMMY = -1
    PACK = 1
    COPY_PIPELINED = 2
    UNPACK = 3
    COPY = 4
    UNPACK_PIPELINED = 5
    PACK_PIPELINED = 6
    PERMUTE_FORWARD = 7
    PERMUTE_BACKWARD = 8
    PERMUTE_BACKWARD_START = 9
    PERMUTE_BACKWARD_END = 10
    PERMUTE_BACKWARD_END_PIPELINED = 11
    PACK_FORWARD = 12
    PACK_BACKWARD = 13
    UNPACK_FORWARD = 15
    UNPACK_FORWARD_PIPELINED = 16
    UNPACK_BACKWARD = 17
    UNPACK_BACKWARD_PIPELINED = 18


class R2RKind(enum.IntEnum):
    """dtfft_r2r_kind_t = FFTW kinds, include/dtfft_config.h.in:66-73"""
    DCT_1 = 3
    DCT_2 = 5
    DCT_3 = 4
    DCT_4 = 6
    DST_1 = 7
    DST_2 = 9
    DST_3 = 8
    DST_4 = 10


class FFTType(enum.IntEnum):
    C2C = 0
    R2C = 1
    R2R = 2


FFT_FORWARD = -1
FFT_BACKWARD = 1


def _ptr(x) -> int | None:
    """Device address of a DeviceArray, an int, or an object exposing data_ptr()."""
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    if hasattr(x, "__cuda_array_interface__"):
        return int(x.__cuda_array_interface__["data"][0])
    raise TypeError(f"cannot take a device address of {type(x)!r}")


class DeviceArray:
    """Owning device buffer (hipMalloc through the library) with numpy transfers."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("dtfft_hip_mem_alloc", C.c_size_t(max(1, self.nbytes)), C.byref(p))
        self.ptr = int(p.value)

    @classmethod
    def from_numpy(cls, a: np.ndarray, nbytes: int | None = None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(max(a.nbytes, nbytes or 0))
        d.upload(a)
        return d

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        call("dtfft_hip_memcpy_h2d", C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p),
             C.c_size_t(a.nbytes))

    def download(self, shape, dtype, offset: int = 0) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        call("dtfft_hip_memcpy_d2h", out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset),
             C.c_size_t(out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().dtfft_hip_mem_free(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    call("dtfft_hip_device_synchronize")


def device_arch() -> str:
    buf = C.create_string_buffer(128)
    call("dtfft_hip_get_device_arch", buf, 128)
    return buf.value.decode()


def set_config(enable_z_slab: bool = False, enable_fused: bool = True, enable_kernel_autotune: bool = False,
               measure_warmup_iters: int = 2, measure_iters: int = 5, stream: int | None = None,
               lds_dma: int = 1, twl: int = 1, blocked: bool = True) -> None:
    """dtfft_set_config for this library (src/dtfft_config.F90).  ``lds_dma``: 0 never, 1 where
    picked (default), 2 the double-buffer kernel wherever allowed, 3 the single-buffer kernel
    wherever allowed -- column-in passes through LDS tiles filled by direct loads.  ``twl``:
    FFT kernels with the twiddle table in LDS (persistent blocks): 0 never, 1 by the default
    rule, 2 wherever the kernel has that form and the LDS budget allows it.  ``blocked``: a row
    pass feeding a column-in pass writes its output chunk-blocked (private layout) so that the
    column pass reads contiguous tiles; False keeps the natural layout."""
    c = Config(int(enable_z_slab), int(enable_fused), int(enable_kernel_autotune), measure_warmup_iters,
               measure_iters, stream, int(lds_dma), int(twl), int(bool(blocked)))
    call("dtfft_hip_set_config", C.byref(c))


def _i32arr(v):
    return (C.c_int32 * len(v))(*[int(x) for x in v])


class Kernel:
    """Kernel plugin handle (dtfft_hip_kernel_*), the drop-in for kernel_device.

    ``dims`` are the local dims, fastest first; ``neighbor_data`` is a sequence of
    (nxx, nyy, nzz, din, dout) rows, one per peer."""

    def __init__(self, kernel_type, dims, base_storage: int, neighbor_data=None,
                 effort: Effort = Effort.ESTIMATE, force_effort: bool = False):
        self.h = C.c_void_p()
        nd = None
        npeers = 0
        if neighbor_data is not None:
            rows = [list(map(int, r)) for r in neighbor_data]
            npeers = len(rows)
            nd = _i32arr([x for r in rows for x in r])
        call("dtfft_hip_kernel_create", int(kernel_type), len(dims), _i32arr(dims), int(base_storage), nd,
             npeers, int(effort), int(bool(force_effort)), C.byref(self.h))

    def execute(self, inp, out, neighbor: int = 0, stream: int | None = None, sync: bool = False) -> None:
        call("dtfft_hip_kernel_execute", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)),
             C.c_void_p(stream), int(neighbor), int(sync))

    def destroy(self) -> None:
        if self.h:
            lib().dtfft_hip_kernel_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class FFTExecutor:
    """FFT executor plugin handle (dtfft_hip_fft_*), the drop-in for abstract_executor."""

    def __init__(self, fft_rank: int, fft_type: FFTType, precision: Precision, idist: int, odist: int,
                 how_many: int, fft_sizes, inembed=None, onembed=None, effort: Effort = Effort.ESTIMATE,
                 stream: int | None = None, r2r_kinds=None):
        self.h = C.c_void_p()
        call("dtfft_hip_fft_create", int(fft_rank), int(fft_type), int(precision), int(idist), int(odist),
             int(how_many), _i32arr(fft_sizes), _i32arr(inembed or fft_sizes), _i32arr(onembed or fft_sizes),
             None if r2r_kinds is None else _i32arr(r2r_kinds), int(effort), C.c_void_p(stream), C.byref(self.h))

    def execute(self, a, b, sign: int) -> None:
        call("dtfft_hip_fft_execute", self.h, C.c_void_p(_ptr(a)), C.c_void_p(_ptr(b)), int(sign))

    def destroy(self) -> None:
        if self.h:
            lib().dtfft_hip_fft_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class _Plan:
    _create = ""

    def __init__(self, dims, precision: Precision = Precision.DOUBLE, effort: Effort = Effort.ESTIMATE,
                 executor: Executor = Executor.NONE):
        """``dims`` in natural Fortran order (X first), as dtFFT's Python PlanC2C."""
        self.dims = [int(d) for d in dims]
        self.h = C.c_void_p()
        rev = list(reversed(self.dims))
        call(self._create, len(rev), _i32arr(rev), int(precision), int(effort), int(executor), C.byref(self.h))

    @property
    def local_sizes(self):
        nd = len(self.dims)
        a, b, c, d = (C.c_int32 * nd)(), (C.c_int32 * nd)(), (C.c_int32 * nd)(), (C.c_int32 * nd)()
        sz = C.c_size_t()
        call("dtfft_hip_get_local_sizes", self.h, a, b, c, d, C.byref(sz))
        # C API reports reversed order; return natural Fortran order like dtFFT's Python
        return (list(a)[::-1], list(b)[::-1], list(c)[::-1], list(d)[::-1], int(sz.value))

    def _get(self, fn) -> int:
        v = C.c_size_t()
        call(fn, self.h, C.byref(v))
        return int(v.value)

    @property
    def alloc_size(self) -> int:
        return self._get("dtfft_hip_get_alloc_size")

    @property
    def element_size(self) -> int:
        return self._get("dtfft_hip_get_element_size")

    @property
    def alloc_bytes(self) -> int:
        return self._get("dtfft_hip_get_alloc_bytes")

    @property
    def aux_bytes(self) -> int:
        return self._get("dtfft_hip_get_aux_bytes")

    @property
    def z_slab_enabled(self) -> bool:
        v = C.c_int32()
        call("dtfft_hip_get_z_slab_enabled", self.h, C.byref(v))
        return bool(v.value)

    def execute(self, inp, out, execute_type: Execute, aux=None) -> None:
        call("dtfft_hip_execute", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)), int(execute_type),
             C.c_void_p(_ptr(aux)))

    def execute_no_alloc(self, inp, out, execute_type: Execute) -> None:
        """execute() with the library's fixed per-device scratch as aux: allocates nothing
        (dtfft_hip_execute_no_alloc); calls that overlap on one GPU race."""
        call("dtfft_hip_execute_no_alloc", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)),
             int(execute_type))

    def execute_cycle_graph(self, inp, out, in2, aux=None) -> None:
        """Forward (inp -> out) then backward (out -> in2), captured once as a hipGraph and
        replayed (dtfft_hip_execute_cycle_graph)."""
        call("dtfft_hip_execute_cycle_graph", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)),
             C.c_void_p(_ptr(aux)), C.c_void_p(_ptr(in2)))

    def execute_stage(self, inp, out, execute_type: Execute, stage: int, aux=None) -> int:
        """Run launch `stage` of the execute sequence only (profiling); returns the
        number of launches in the sequence (stage < 0 runs nothing)."""
        n = C.c_int32()
        call("dtfft_hip_plan_execute_stage", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)),
             int(execute_type), C.c_void_p(_ptr(aux)), int(stage), C.byref(n))
        return int(n.value)

    def transpose(self, inp, out, transpose_type: Transpose, aux=None) -> None:
        call("dtfft_hip_transpose", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)), int(transpose_type),
             C.c_void_p(_ptr(aux)))

    def mem_alloc(self, nbytes: int | None = None) -> DeviceArray:
        return DeviceArray(nbytes or self.alloc_bytes)

    def report(self) -> None:
        call("dtfft_hip_report", self.h)

    def destroy(self) -> None:
        if self.h:
            lib().dtfft_hip_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PlanC2C(_Plan):
    """Complex-to-complex plan (dtfft_create_plan_c2c for one rank)."""
    _create = "dtfft_hip_create_plan_c2c"


class PlanR2C(_Plan):
    """Real-to-complex plan (dtfft_create_plan_r2c for one rank)."""
    _create = "dtfft_hip_create_plan_r2c"


class PlanR2R(_Plan):
    """Real-to-real plan (dtfft_create_plan_r2r for one rank).  ``kinds`` in natural
    Fortran order (X first), like ``dims``; None for a transpose-only plan."""
    _create = "dtfft_hip_create_plan_r2r"

    def __init__(self, dims, kinds=None, precision: Precision = Precision.DOUBLE, effort: Effort = Effort.ESTIMATE,
                 executor: Executor = Executor.NONE):
        self.dims = [int(d) for d in dims]
        self.kinds = None if kinds is None else [int(k) for k in kinds]
        self.h = C.c_void_p()
        rev = list(reversed(self.dims))
        kv = None if kinds is None else _i32arr(list(reversed(self.kinds)))
        call(self._create, len(rev), _i32arr(rev), kv, int(precision), int(effort), int(executor),
             C.byref(self.h))


class UnpackFFT:
    """Receive side of a P > 1 transpose fused with the FFT along the aligned axis
    (dtfft_hip_unpack_fft_*).  ``dims`` are the receive pencil's local dims (aligned
    axis first), ``recv_rows`` the unpack neighbour data."""

    def __init__(self, dims, recv_rows, precision: Precision, effort: Effort = Effort.ESTIMATE,
                 stream: int | None = None):
        self.h = C.c_void_p()
        flat = [int(x) for r in recv_rows for x in r]
        call("dtfft_hip_unpack_fft_create", len(dims), _i32arr(dims), _i32arr(flat), len(recv_rows),
             int(precision), int(effort), C.c_void_p(stream), C.byref(self.h))

    def execute(self, recv, out, sign: int = FFT_FORWARD) -> None:
        call("dtfft_hip_unpack_fft_execute", self.h, C.c_void_p(_ptr(recv)), C.c_void_p(_ptr(out)), int(sign))

    def destroy(self) -> None:
        if self.h:
            lib().dtfft_hip_unpack_fft_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class ExchangeShare:
    """One GPU's share of a forward transpose on P peers (equal splits), with a loopback
    copy in place of the exchange: PACK_FORWARD per peer into contiguous per-peer send
    chunks, a copy kernel moving the send buffer into the receive buffer, then either UNPACK +
    FFT along the aligned axis (``fused=False``) or the fused unpack+FFT (``fused=True``).
    Pencil A = [a, b_me, c] -> B = [b, c, a_me]."""

    def __init__(self, a: int, b: int, c: int, peers: int, me: int, precision: Precision, fused: bool = True,
                 effort: Effort = Effort.ESTIMATE):
        from . import pencil
        if a % peers or b % peers:
            raise ValueError("the loopback share needs equal splits")
        self.a, self.b, self.c, self.P, self.me = a, b, c, peers, me
        self.es = 16 if precision == Precision.DOUBLE else 8
        send, recv, sc, rc = pencil.forward_exchange(a, b, c, peers, me)
        self.send_rows, self.recv_rows = send, recv
        self.bytes = sum(sc) * self.es
        bme, ame = b // peers, a // peers
        self.a_dims, self.b_dims = (a, bme, c), (b, c, ame)
        self.pack = Kernel(KernelType.PACK_FORWARD, self.a_dims, self.es, send, effort)
        # the loopback in place of the exchange: the library's copy kernel (16-byte lanes) over the
        # whole send buffer, one box of sum(counts) elements
        n_send = sum(sc)
        self.loopback = Kernel(KernelType.PACK, [n_send, 1], self.es, [[n_send, 1, 1, 0, 0]], effort)
        self.fused = fused
        if fused:
            self.unpack_fft = UnpackFFT(self.b_dims, recv, precision, effort)
        else:
            self.unpack = Kernel(KernelType.UNPACK, self.b_dims, self.es, recv, effort)
            self.fft = FFTExecutor(1, FFTType.C2C, precision, b, b, c * ame, [b], effort=effort)

    def run(self, A, send, recv, B, sign: int = FFT_FORWARD, tmp=None) -> None:
        for i in range(1, self.P + 1):
            self.pack.execute(A, send, neighbor=i)
        self.loopback.execute(send, recv)
        if self.fused:
            self.unpack_fft.execute(recv, B, sign)
        else:
            self.unpack.execute(recv, tmp)
            self.fft.execute(tmp, B, sign)


class PencilTranspose:
    """One rank's transpose of a P > 1 plan on the GPU: the pack and unpack kernels that
    reshape_handle_generic%create picks, with their neighbour data
    (src/dtfft_reshape_handle_generic.F90:174-693, restated in pencil.TransposeShare), run in
    the pack -> exchange -> unpack order of its execute (:695-759).  Every transpose type,
    both directions (the 3D backward ones as the two-step PERMUTE_BACKWARD_START /
    PERMUTE_BACKWARD_END pair), uneven splits, modes "mpi" / "pipelined" / "fused" /
    "unpack".  The exchange is the caller's: on one GPU, a device copy of each peer chunk
    (send_counts / send_displs to recv_displs of the peer)."""

    def __init__(self, transpose_type: int, dims, comm_size: int, me: int, base_storage: int, mode: str = "mpi",
                 effort: Effort = Effort.ESTIMATE, fixed=(None, 0)):
        from . import pencil
        self.share = sh = pencil.TransposeShare(transpose_type, dims, comm_size, me, mode, fixed)
        self.es = int(base_storage)
        self.pack = Kernel(sh.pack_kernel, sh.in_counts, self.es, sh.send_rows, effort)
        self.unpack = Kernel(sh.unpack_kernel, sh.out_counts, self.es, sh.recv_rows, effort)

    @property
    def send_bytes(self) -> int:
        return sum(self.share.send_counts) * self.es

    @property
    def recv_bytes(self) -> int:
        return sum(self.share.recv_counts) * self.es

    def pack_into(self, inp, send, stream: int | None = None) -> None:
        self.share.run_pack(lambda i: self.pack.execute(inp, send, neighbor=i, stream=stream))

    def unpack_from(self, recv, out, stream: int | None = None) -> None:
        self.share.run_unpack(lambda i: self.unpack.execute(recv, out, neighbor=i, stream=stream))


class PencilReshape:
    """One rank's bricks <-> pencils reshape on the GPU (src/dtfft_reshape_plan.F90:98-260 with
    the reshape branches of reshape_handle_generic, restated in pencil.ReshapeShare): PACK /
    UNPACK kernels with their neighbour data, or none where the layout is pack-free /
    unpack-free, in the pack -> exchange -> unpack order.  The exchange is the caller's."""

    def __init__(self, reshape_type: int, dims, comm_size: int, me: int, base_storage: int, mode: str = "mpi",
                 split=None, effort: Effort = Effort.ESTIMATE):
        from . import pencil
        self.share = sh = pencil.ReshapeShare(reshape_type, dims, comm_size, me, mode, split)
        self.es = int(base_storage)
        self.pack = self.unpack = None
        if sh.pack_kernel != pencil.KERNEL_DUMMY:
            self.pack = Kernel(sh.pack_kernel, sh.in_counts, self.es, sh.send_rows, effort)
        if sh.unpack_kernel != pencil.KERNEL_DUMMY:
            self.unpack = Kernel(sh.unpack_kernel, sh.out_counts, self.es, sh.recv_rows, effort)

    @property
    def send_bytes(self) -> int:
        return sum(self.share.send_counts) * self.es

    @property
    def recv_bytes(self) -> int:
        return sum(self.share.recv_counts) * self.es

    def pack_into(self, inp, send, stream: int | None = None) -> None:
        """With a pack-free layout nothing runs: the input is the send buffer."""
        if self.pack is not None:
            self.share.run_pack(lambda i: self.pack.execute(inp, send, neighbor=i, stream=stream))

    def unpack_from(self, recv, out, stream: int | None = None) -> None:
        """With an unpack-free layout nothing runs: the receive buffer is the output."""
        if self.unpack is not None:
            self.share.run_unpack(lambda i: self.unpack.execute(recv, out, neighbor=i, stream=stream))


class Cycle:
    """Batched fwd+bwd cycle over many 3D C2C grids (dims per grid in Fortran order)."""

    def __init__(self, grids, precision: Precision = Precision.SINGLE, effort: Effort = Effort.ESTIMATE,
                 stream: int | None = None):
        self.grids = [[int(d) for d in g] for g in grids]
        flat = [d for g in self.grids for d in reversed(g)]
        self.h = C.c_void_p()
        call("dtfft_hip_cycle_create", len(self.grids), _i32arr(flat), int(precision), int(effort),
             C.c_void_p(stream), C.byref(self.h))

    def layout(self):
        offs = (C.c_int64 * len(self.grids))()
        sz = C.c_size_t()
        call("dtfft_hip_cycle_get_layout", self.h, offs, C.byref(sz))
        return list(offs), int(sz.value)

    def execute(self, inp, out, aux, execute_type: Execute) -> None:
        call("dtfft_hip_cycle_execute", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)),
             C.c_void_p(_ptr(aux)), int(execute_type))

    def execute_graph(self, inp, out, aux, in2) -> None:
        call("dtfft_hip_cycle_execute_graph", self.h, C.c_void_p(_ptr(inp)), C.c_void_p(_ptr(out)),
             C.c_void_p(_ptr(aux)), C.c_void_p(_ptr(in2)))

    def destroy(self) -> None:
        if self.h:
            lib().dtfft_hip_cycle_destroy(C.byref(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
