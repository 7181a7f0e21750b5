"""GPU parity of the batched cycle (many uneven grids, one persistent launch per stage,
fwd+bwd captured as a hipGraph) against the oracle, grid by grid."""
import numpy as np
import pytest

import oracle.plan as OP

pytestmark = pytest.mark.gpu

D = pytest.importorskip("dtfft_amd")

SIZES = [48, 60, 64, 84, 96, 105, 128, 160, 192, 210, 256, 320, 384]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def pack(grids, offs, total, rng, dtype):
    buf = np.zeros(total, dtype=dtype)
    xs = []
    for g, o in zip(grids, offs):
        nx, ny, nz = g
        n = nx * ny * nz
        x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(dtype)
        buf[o:o + n] = x
        xs.append(x.reshape(nz, ny, nx))
    return buf, xs


@pytest.mark.parametrize("prec", [D.Precision.SINGLE, D.Precision.DOUBLE])
def test_cycle_small_against_oracle(prec):
    grids = [[12, 10, 6], [16, 16, 16], [60, 21, 10], [7, 13, 5], [48, 30, 20]]
    dt = np.complex64 if prec == D.Precision.SINGLE else np.complex128
    tol = 1e-6 if prec == D.Precision.SINGLE else 1e-12
    cy = D.Cycle(grids, prec)
    offs, total = cy.layout()
    rng = np.random.default_rng(7)
    buf, xs = pack(grids, offs, total, rng, dt)
    es = np.dtype(dt).itemsize
    a = D.DeviceArray.from_numpy(buf)
    b = D.DeviceArray(total * es)
    aux = D.DeviceArray(total * es)
    c = D.DeviceArray(total * es)
    cy.execute(a, b, aux, D.Execute.FORWARD)
    out = b.download(total, dt)
    for g, o, x in zip(grids, offs, xs):
        nx, ny, nz = g
        z = out[o:o + x.size].reshape(ny, nx, nz).astype(np.complex128)
        assert rel(z, OP.execute_generic(x.astype(np.complex128), True)) < tol, g
    cy.execute(b, c, aux, D.Execute.BACKWARD)
    back = c.download(total, dt)
    for g, o, x in zip(grids, offs, xs):
        y = back[o:o + x.size].astype(np.complex128) / x.size
        assert rel(y, x.ravel().astype(np.complex128)) < tol * 10, g


def test_cycle_graph_replays_match_eager():
    """The hipGraph of fwd+bwd is replayed several times with eager work and syncs in
    between; every replay must equal the eager result bit for bit."""
    rng = np.random.default_rng(11)
    grids = [list(map(int, rng.choice(SIZES[:6], 3))) for _ in range(6)]
    cy = D.Cycle(grids, D.Precision.SINGLE)
    offs, total = cy.layout()
    buf, xs = pack(grids, offs, total, rng, np.complex64)
    a = D.DeviceArray.from_numpy(buf)
    b, aux, c = D.DeviceArray(total * 8), D.DeviceArray(total * 8), D.DeviceArray(total * 8)
    cy.execute(a, b, aux, D.Execute.FORWARD)
    cy.execute(b, c, aux, D.Execute.BACKWARD)
    eager_fwd = b.download(total, np.complex64)
    eager_back = c.download(total, np.complex64)
    for rep in range(4):
        b.upload(np.zeros(total, np.complex64))
        c.upload(np.zeros(total, np.complex64))
        cy.execute_graph(a, b, aux, c)
        D.synchronize()
        assert b.download(total, np.complex64).tobytes() == eager_fwd.tobytes(), rep
        assert c.download(total, np.complex64).tobytes() == eager_back.tobytes(), rep
        cy.execute(a, b, aux, D.Execute.FORWARD)  # eager work between replays


def test_cycle_64_uneven_round_trip():
    """Batched-uneven workload: 64 grids, axis lengths drawn from the seed out of the
    radix-2/3/5/7 set; bwd(fwd(x))/N == x within 1e-6 for every grid (FP32)."""
    rng = np.random.default_rng(20250104)
    grids = [list(map(int, rng.choice(SIZES, 3))) for _ in range(64)]
    cy = D.Cycle(grids, D.Precision.SINGLE)
    offs, total = cy.layout()
    buf, xs = pack(grids, offs, total, rng, np.complex64)
    a = D.DeviceArray.from_numpy(buf)
    b, aux, c = D.DeviceArray(total * 8), D.DeviceArray(total * 8), D.DeviceArray(total * 8)
    cy.execute_graph(a, b, aux, c)
    D.synchronize()
    back = c.download(total, np.complex64)
    worst = 0.0
    for g, o, x in zip(grids, offs, xs):
        y = back[o:o + x.size].astype(np.complex128) / x.size
        worst = max(worst, rel(y, x.ravel().astype(np.complex128)))
    assert worst < 1e-6, worst
    # spot-check three grids' forward output against the oracle
    out = b.download(total, np.complex64)
    for i in (0, 17, 63):
        nx, ny, nz = grids[i]
        z = out[offs[i]:offs[i] + xs[i].size].reshape(ny, nx, nz).astype(np.complex128)
        assert rel(z, OP.execute_generic(xs[i].astype(np.complex128), True)) < 1e-6, grids[i]


@pytest.mark.parametrize("effort", [D.Effort.MEASURE, D.Effort.EXHAUSTIVE])
def test_tuned_cycle_against_oracle(tmp_path, effort):
    """At effort >= MEASURE the batched cycle times candidate tile shapes and radix schedules
    for every length group and keeps the fastest (tuning records "cyc|..."); the tuned cycle
    must match the oracle grid by grid, and a second cycle reads the records back."""
    import os
    L = D.lib()
    default_dir = os.environ.get("DTFFT_HIP_CACHE_DIR") or os.path.join(os.path.dirname(D._lib.LIB_PATH), "kcache")
    assert L.dtfft_hip_set_cache_dir(str(tmp_path).encode()) == 0
    try:
        grids = [[48, 30, 20], [64, 48, 60], [105, 32, 16], [384, 16, 12], [320, 20, 16]]
        for rnd in range(2):
            cy = D.Cycle(grids, D.Precision.SINGLE, effort)
            offs, total = cy.layout()
            rng = np.random.default_rng(9 + rnd)
            buf, xs = pack(grids, offs, total, rng, np.complex64)
            a = D.DeviceArray.from_numpy(buf)
            b, aux, c = D.DeviceArray(total * 8), D.DeviceArray(total * 8), D.DeviceArray(total * 8)
            cy.execute(a, b, aux, D.Execute.FORWARD)
            out = b.download(total, np.complex64)
            for g, o, x in zip(grids, offs, xs):
                nx, ny, nz = g
                z = out[o:o + x.size].reshape(ny, nx, nz).astype(np.complex128)
                assert rel(z, OP.execute_generic(x.astype(np.complex128), True)) < 1e-6, g
            cy.execute(b, c, aux, D.Execute.BACKWARD)
            back = c.download(total, np.complex64)
            for g, o, x in zip(grids, offs, xs):
                assert rel(back[o:o + x.size].astype(np.complex128) / x.size, x.ravel()) < 1e-5, g
        recs = [ln for ln in (tmp_path / "tuning.txt").read_text().splitlines() if ln.startswith("cyc|")]
        assert len(recs) >= 6
    finally:
        L.dtfft_hip_set_cache_dir(default_dir.encode())
