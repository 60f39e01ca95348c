"""GPU Kaiser-Bessel (width 6) NUFFT variant against exact transforms.

The reference has only the Gaussian window, so this variant is checked
against mathematics, not the FP32 mirror: the forward and adjoint against the
FP64 non-uniform DFT (the window's error is ~5e-6 relative at sigma = 2), the
device operators against each other (adjointness), and whole solves against
the FP64 oracle, whose own error against the exact solve is far smaller.
"""
import numpy as np
import pytest

from feilab_amd import phantom

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


GEOMS = [(96, 96, phantom.tilt_series(60, 3)), (128, 64, phantom.tilt_series(57, 4)),
         (200, 200, phantom.tilt_series(60, 5))]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "%dx%d" % (g[0], g[1]))
def test_kb_forward_adjoint_vs_fp64_ndft(icon, oracle_mod, geom):
    N0, N1, tilt = geom
    print("kb ndft %dx%d" % (N0, N1), flush=True)
    rng = np.random.default_rng(N0 + N1)
    plan = icon.Plan(N0, tilt, thickness=N1, window="kb6")
    assert plan.info.window == 1 and plan.info.m == 3
    nodes = oracle_mod.Plan(N0, N1, tilt, 64).tables()["nodes"]
    img = rng.standard_normal((N0, N1))
    assert rel(plan.forward(img.astype(np.float32)), oracle_mod.ndft_forward(nodes, img.astype(np.complex128))) < 2e-5
    f = rng.standard_normal(len(nodes)) + 1j * rng.standard_normal(len(nodes))
    assert rel(plan.adjoint(f.astype(np.complex64)), oracle_mod.ndft_adjoint(nodes, f, N0, N1)) < 2e-5


def test_kb_adjointness_and_window_gain(icon, oracle_mod):
    """<A x, y> = <x, A^H y> for the device operators; the Gaussian plan of the
    same geometry is the more accurate one, the KB plan still within 2e-5."""
    N0, N1, tilt = GEOMS[0]
    rng = np.random.default_rng(9)
    kb = icon.Plan(N0, tilt, thickness=N1, window="kb6")
    gs = icon.Plan(N0, tilt, thickness=N1)
    nodes = oracle_mod.Plan(N0, N1, tilt, 64).tables()["nodes"]
    x = rng.standard_normal((N0, N1)).astype(np.float32)
    y = (rng.standard_normal(len(nodes)) + 1j * rng.standard_normal(len(nodes))).astype(np.complex64)
    lhs = np.vdot(y.astype(np.complex128), kb.forward(x).astype(np.complex128))
    rhs = np.vdot(kb.adjoint(y).astype(np.complex128), x.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    exact = oracle_mod.ndft_forward(nodes, x.astype(np.complex128))
    e_kb, e_gs = rel(kb.forward(x), exact), rel(gs.forward(x), exact)
    print("forward error vs exact: kb6 %.2e gauss %.2e" % (e_kb, e_gs), flush=True)
    assert e_gs < e_kb < 2e-5


def test_kb_sino_to_b_vs_fp64(icon, oracle_mod):
    cfg = phantom.config1()
    plan = icon.Plan(256, cfg["tilt"], max_slices=1, window="kb6")
    orc = oracle_mod.Plan(256, 256, cfg["tilt"], 64)
    assert rel(plan.sino_to_b(cfg["sino"][None])[0], orc.sino_to_b(cfg["sino"])) < 2e-5


@pytest.mark.parametrize("data_type,thr", [(2, 0.0), (1, 0.03)])
def test_kb_solve_vs_fp64_oracle(icon, oracle_mod, data_type, thr):
    print("kb solve", data_type, thr, flush=True)
    rng = np.random.default_rng(11)
    N = 96
    tilt = phantom.tilt_series(60, 3)
    sino = phantom.add_noise(phantom.project_2d(phantom.blobs_2d(N, 12, rng), N, tilt), 0.5, rng)
    plan = icon.Plan(N, tilt, max_slices=1, window="kb6")
    m, rp, res, _ = plan.solve(sino, 5, 30, 5, data_type, thr, trace=True, reproject=True)
    orc = oracle_mod.Plan(N, N, tilt, 64)
    mo, reso, _ = orc.solve(sino, 5, 30, 5, data_type, thr, trace=True)
    err = rel(m, mo)
    print("kb6 solve vs FP64 oracle: rel L2 %.2e, final residual %.6f vs %.6f" % (err, res[-1], reso[-1]), flush=True)
    assert err < 2e-4
    assert abs(res[-1] - reso[-1]) < 1e-4 * reso[-1]
    assert rel(rp, orc.reproject(mo)) < 2e-4
