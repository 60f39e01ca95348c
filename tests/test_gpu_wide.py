"""The widest grids and the 16-column tile, against the oracle.

* Widths and thicknesses 513..1024 give the oversampled grid n = 2048
  (n = 2 nextpow2(N), MainCode.cu:66-71): the column kernels at n0 = 2048
  (one 1024-thread block per CU, interpolation rows read element by element
  because the LDS row stride C + 1 is odd) and the row kernels at n1 = 2048.
  Forward and adjoint NUFFT against the FP32 mirror and the FP64 oracle, and
  a short solve against the FP32 mirror.
* C = 16 at n0 = 1024 (also an odd LDS row stride) gives the C = 8 solve to
  rounding.
"""
import numpy as np
import pytest

from feilab_amd import phantom

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.parametrize("N0,N1", [(600, 600), (1024, 520)])
def test_nufft_at_grid_2048(icon, oracle_mod, N0, N1):
    tilt = phantom.tilt_series(60, 10)
    print("nufft at n = 2048: %d x %d, %d tilts" % (N0, N1, len(tilt)), flush=True)
    rng = np.random.default_rng(N0)
    plan = icon.Plan(N0, tilt, thickness=N1, max_slices=2)
    assert plan.n0 == 2048 and plan.n1 == 2048
    o32 = oracle_mod.Plan(N0, N1, tilt, 32)
    o64 = oracle_mod.Plan(N0, N1, tilt, 64)
    img = rng.standard_normal((2, N0, N1)).astype(np.float32)
    f = plan.forward(img)
    for s in range(2):
        f64 = o64.forward(img[s].astype(np.complex128))
        e32, e64 = rel(f[s], o32.forward(img[s].astype(np.complex128))), rel(f[s], f64)
        print("  forward slice %d: vs FP32 mirror %.2e, vs FP64 %.2e" % (s, e32, e64), flush=True)
        assert e32 < 2e-5 and e64 < 2e-5
    g = (rng.standard_normal((2, plan.M)) + 1j * rng.standard_normal((2, plan.M))).astype(np.complex64)
    a = plan.adjoint(g)
    for s in range(2):
        e32, e64 = rel(a[s], o32.adjoint(g[s])), rel(a[s], o64.adjoint(g[s].astype(np.complex128)))
        print("  adjoint slice %d: vs FP32 mirror %.2e, vs FP64 %.2e" % (s, e32, e64), flush=True)
        assert e32 < 2e-5 and e64 < 2e-5


def test_solve_at_grid_2048(icon, oracle_mod):
    N = 600
    tilt = phantom.tilt_series(60, 6)
    print("solve at n = 2048: %d^2, %d tilts, 2 slices, 12 iterations" % (N, len(tilt)), flush=True)
    rng = np.random.default_rng(77)
    b = phantom.blobs_3d(N, 2, N, 30, rng)
    sinos = phantom.add_noise(phantom.project_3d(b, N, np.arange(2), tilt), 0.2, rng)
    plan = icon.Plan(N, tilt, max_slices=2)
    m, res, _ = plan.solve(sinos, 3, 6, 3, 2, 0.0, trace=True)
    o32 = oracle_mod.Plan(N, N, tilt, 32)
    mo, reso, _ = o32.solve_batch(sinos, 3, 6, 3, 2, 0.0, threads=2, trace=True)
    for s in range(2):
        e = rel(m[s], mo[s])
        print("  slice %d: rel L2 vs FP32 mirror %.2e, R %.6f vs %.6f" % (s, e, res[s][-1], reso[s][-1]), flush=True)
        assert e < 1e-4 and abs(res[s][-1] - reso[s][-1]) < 1e-4 * reso[s][-1]


def test_col_block_16_at_grid_1024(icon):
    cfg = phantom.config2(nslices=2, nblobs=40)
    print("C = 16 vs C = 8 at n0 = 1024, 2 slices, 12 iterations", flush=True)
    plan = icon.Plan(512, cfg["tilt"], max_slices=2)
    m8, r8, _ = plan.solve(cfg["sino"], 3, 6, 3, 2, 0.0, trace=True)
    plan.set_col_block(16)
    assert plan.info.col_block == 16
    m16, r16, _ = plan.solve(cfg["sino"], 3, 6, 3, 2, 0.0, trace=True)
    e = rel(m16, m8)
    print("  rel L2 C=16 vs C=8: %.2e" % e, flush=True)
    assert e < 1e-5 and np.allclose(r16, r8, rtol=1e-5)
