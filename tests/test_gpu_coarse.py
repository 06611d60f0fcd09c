"""The learned two-grid preconditioner of the flexible CG (deck keyword
tl_fcg_coarse): its coarse operator and restriction against a numpy Galerkin
construction, and what the trained network gains over the sweeps alone.

With piecewise-constant prolongation P over B x B blocks, the coarse
operator P^T A P / B^2 has unit diagonal plus face coefficients: the sums
of the fine face coefficients across each coarse face over B^2.  No
counterpart in the reference.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import decks

pytestmark = pytest.mark.gpu

HALO = 2


def fine_operator(kx, ky):
    """A = I + sum over interior faces k (e_i - e_j) (no flux across the mesh
    boundary), kx[y, x] the west and ky[y, x] the south face of cell (x, y)."""
    ny, nx = kx.shape
    idx = np.arange(nx * ny).reshape(ny, nx)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [np.ones(nx * ny)]
    for (a, b, k) in ((idx[:, 1:], idx[:, :-1], kx[:, 1:]), (idx[1:, :], idx[:-1, :], ky[1:, :])):
        a, b, k = a.ravel(), b.ravel(), k.ravel()
        rows += [a, b, a, b]
        cols += [a, b, b, a]
        vals += [k, k, -k, -k]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(nx * ny, nx * ny))


def prolongation(nx, ny, B):
    ncx = nx // B
    blk = (np.arange(ny)[:, None] // B) * ncx + (np.arange(nx)[None, :] // B)
    return sp.csr_matrix((np.ones(nx * ny), (np.arange(nx * ny), blk.ravel())),
                         shape=(nx * ny, ncx * (ny // B)))


@pytest.mark.parametrize("nx,ny,nc", [(128, 64, 16), (96, 96, 24), (64, 128, 8)])
def test_coarse_operator_and_restriction(hip, nx, ny, nc):
    g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
    rng = np.random.default_rng(nx + 7 * ny + nc)
    for name, lo, hi in (("kx", 0.1, 50.0), ("ky", 0.1, 50.0), ("density", 0.1, 100.0), ("r", -1, 1), ("z", -1, 1)):
        g.set(name, rng.uniform(lo, hi, size=g.shape))
    B = nx // nc
    ncy = ny // B
    n = nc * ncy
    kxc, kyc, fc = (np.zeros(n, np.float32) for _ in range(3))
    rc = np.zeros(n)
    P32 = ctypes.POINTER(ctypes.c_float)
    P64 = ctypes.POINTER(ctypes.c_double)
    for with_z in (1, 0):
        rc_k = np.zeros(n)
        assert g.L.tl_coarse_probe(g.c, nc, with_z, kxc.ctypes.data_as(P32), kyc.ctypes.data_as(P32),
                                   fc.ctypes.data_as(P32), rc_k.ctypes.data_as(P64)) == 0, g.L.tl_last_error()
        if with_z:
            rc = rc_k
        else:
            rc0 = rc_k
    inner = (0, slice(HALO, HALO + ny), slice(HALO, HALO + nx))
    kx, ky = g.get("kx")[inner], g.get("ky")[inner]
    A = fine_operator(kx, ky)
    P = prolongation(nx, ny, B)
    Ac = (P.T @ A @ P).toarray() / B**2
    cidx = np.arange(n).reshape(ncy, nc)
    west = np.zeros((ncy, nc))
    south = np.zeros((ncy, nc))
    west[:, 1:] = -Ac[cidx[:, 1:].ravel(), cidx[:, :-1].ravel()].reshape(ncy, nc - 1)
    south[1:, :] = -Ac[cidx[1:, :].ravel(), cidx[:-1, :].ravel()].reshape(ncy - 1, nc)
    np.testing.assert_allclose(kxc.reshape(ncy, nc), west, rtol=2e-6, atol=0)
    np.testing.assert_allclose(kyc.reshape(ncy, nc), south, rtol=2e-6, atol=0)
    # unit diagonal plus the four faces
    faces = west + south
    faces[:, :-1] += west[:, 1:]
    faces[:-1, :] += south[1:, :]
    np.testing.assert_allclose(np.diag(Ac).reshape(ncy, nc), 1.0 + faces, rtol=1e-12)
    feat = (np.log10(g.get("density")[inner]) - 0.5) / 1.5
    np.testing.assert_allclose(fc, (P.T @ feat.ravel()) / B**2, rtol=2e-6, atol=1e-6)
    r, z = g.get("r")[inner].ravel(), g.get("z")[inner].ravel()
    want = P.T @ (r - A @ z) / B**2
    scale = P.T @ (np.abs(r) + np.abs(A) @ np.abs(z)) / B**2
    assert np.all(np.abs(rc - want) <= 1e-12 * scale)
    want0 = P.T @ r / B**2
    assert np.all(np.abs(rc0 - want0) <= 1e-12 * (P.T @ np.abs(r)) / B**2)
    g.close()


def test_coarse_probe_refuses_bad_grids(hip):
    g = hip.ExtChunk(hip.lib(2), 2, 96, 40)
    for nc in (7, 10, 96, 48):  # not a multiple of 8 / not dividing / B < 2 / y blocks not a multiple of 8
        assert g.L.tl_coarse_probe(g.c, nc, 1, None, None, None, None) == -1
        assert g.L.tl_last_error()
        g.L.tl_clear_error()
    g.close()


def test_learned_two_grid_gains(hip):
    """Trained on whole-coarse-grid pairs from FP64 PPCG solves of 256^2
    layouts (32^2 coarse cells, dt doubled so the coarse operator matches),
    the network as the coarse solve cuts the flexible CG's iterations on
    held-out 512^2 layouts against the same flexible CG with the sweeps
    alone; both end below eps on the true residual, with the FP64 PPCG
    solution's totals."""
    learner = hip.Learner(seed=5)
    learner.loss_smoothing(0, 0.8)
    for k in range(4):
        learner.add_deck_coarse(decks.random_shapes_2d(256, 1000 + k, steps=1, dt=0.008), snapshots=16, nc=32)
    losses = []
    for chunk in range(6):
        losses += list(learner.train(100, batch=64, lr=2e-3 * (0.05 + 0.95 * (1 - chunk / 6))))
    print(f"two-grid training: {learner.pairs} pairs, loss {np.mean(losses[:20]):.3f} -> {np.mean(losses[-20:]):.3f}",
          flush=True)
    assert np.mean(losses[-20:]) < 0.5 * np.mean(losses[:20])
    learner.use(True)
    try:
        gains = []
        for seed in (0, 1):
            deck = decks.random_shapes_2d(512, seed, steps=1, extra=("tl_check_result",))
            # the fixed sweep weight the network's loss is measured after
            fdeck = deck.replace("tl_use_ppcg", "tl_use_fcg_learned\ntl_fcg_omega=0.8")
            ref = hip.run_deck(deck, dim=2)["steps"][0]
            sweeps = hip.run_deck(fdeck.replace("tl_use_fcg_learned", "tl_use_fcg_learned\ntl_fcg_no_network"),
                                  dim=2)["steps"][0]
            coarse = hip.run_deck(fdeck.replace("tl_use_fcg_learned", "tl_use_fcg_learned\ntl_fcg_coarse"),
                                  dim=2)["steps"][0]
            print(f"seed {seed}: sweeps {sweeps['fcg_iters']} iterations, two-grid {coarse['fcg_iters']}", flush=True)
            for s in (sweeps, coarse):
                assert not s["fcg_fallback"]
                assert s["exact_error"] < 1e-10
                assert abs(s["temp"] - ref["temp"]) <= 1e-8 * abs(ref["temp"])
            gains.append(sweeps["fcg_iters"] / coarse["fcg_iters"])
        assert min(gains) > 1.2, gains
    finally:
        learner.use(False)
