"""Device batched SVD of the two-site split (one-sided Jacobi, FP64) against
LAPACK and against matrices of known singular values, and a DMRG sweep with
the device split against the host split."""
import numpy as np
import pytest

from vmps_amd import dmrg

pytestmark = pytest.mark.gpu


def _known(rng, m, n, sv):
    """m x n matrix with singular values sv (orthogonal factors from QR)."""
    k = len(sv)
    U, _ = np.linalg.qr(rng.standard_normal((m, k)))
    V, _ = np.linalg.qr(rng.standard_normal((n, k)))
    return (U * sv) @ V.T


def test_block_svd_shapes_and_accuracy():
    rng = np.random.default_rng(1)
    shapes = [(1, 1), (5, 1), (1, 7), (13, 13), (40, 17), (17, 40), (128, 96), (200, 200), (3, 300), (0, 4), (4, 0)]
    mats = [rng.standard_normal(s) for s in shapes]
    mats.append(_known(rng, 60, 50, np.logspace(0, -12, 50)))        # graded spectrum
    mats.append(_known(rng, 70, 70, np.r_[np.ones(20), np.zeros(50)]))  # rank 20, degenerate
    out = dmrg.block_svd_device(mats)
    for A, (U, S, Vt) in zip(mats, out):
        k = min(A.shape)
        print("shape", A.shape, flush=True)
        assert U.shape == (A.shape[0], k) and S.shape == (k,) and Vt.shape == (k, A.shape[1])
        if k == 0:
            continue
        assert np.all(np.diff(S) <= 0) and S[-1] >= 0
        ref = np.linalg.svd(A, compute_uv=False)
        scale = max(ref[0], 1e-300)
        assert np.max(np.abs(S - ref)) <= 1e-13 * scale
        assert np.linalg.norm(A - (U * S) @ Vt) <= 1e-13 * max(np.linalg.norm(A), 1e-300)
        keep = S > 1e-10 * scale
        assert np.allclose(U[:, keep].T @ U[:, keep], np.eye(keep.sum()), atol=1e-12)
        assert np.allclose(Vt @ Vt.T, np.eye(k), atol=1e-12)
    # the graded spectrum: every singular value down to 1e-12 to 1e-13 of the largest
    S = out[len(shapes)][1]
    assert np.max(np.abs(S - np.logspace(0, -12, 50))) <= 1e-14


@pytest.mark.parametrize("model,L,M", [("HeisenbergU1", 20, 160), ("HubbardSU2xU1", 12, 128)])
def test_dmrg_device_split_matches_host_split(model, L, M):
    """Sweeps with the split SVDs on the device reach the energy of the same
    sweeps with host LAPACK."""
    params = {"t": 1.0, "U": 8.0} if "Hubbard" in model else {"J": 1.0}
    be_dev = dmrg.GpuBackend(device_svd_min=1)
    r_dev = dmrg.run_dmrg(model, L, M, 2, backend=be_dev, params=params)
    be_host = dmrg.GpuBackend(device_svd_min=0)
    r_host = dmrg.run_dmrg(model, L, M, 2, backend=be_host, params=params)
    print(model, "device", r_dev["energy"], "host", r_host["energy"], "device splits", be_dev.device_svds, flush=True)
    assert be_dev.device_svds > 0 and be_host.device_svds == 0
    assert abs(r_dev["energy"] - r_host["energy"]) <= 1e-9 * abs(r_host["energy"])


def test_block_eigh_split_against_lapack():
    """Density-matrix factorisation on the device: lambda = sigma^2 (LAPACK),
    W orthonormal, W C (side 0) or C W^T (side 1) reproduces the matrix."""
    rng = np.random.default_rng(2)
    shapes = [(1, 1), (5, 1), (1, 7), (40, 17), (17, 40), (300, 260), (0, 4), (4, 0)]
    mats = [rng.standard_normal(s) for s in shapes]
    mats.append(_known(rng, 60, 50, np.logspace(0, -6, 50)))
    for side in (0, 1):
        out = dmrg.block_eigh_device(mats, side)
        for A, (lam, W, C) in zip(mats, out):
            print("side", side, "shape", A.shape, flush=True)
            k = A.shape[0] if side == 0 else A.shape[1]
            assert lam.shape == (k,) and W.shape == (k, k) and C.shape == A.shape
            if k == 0 or A.size == 0:
                continue
            sv = np.linalg.svd(A, compute_uv=False)
            ref = np.zeros(k)
            ref[:len(sv)] = sv ** 2
            assert np.all(np.diff(lam) <= 1e-14 * lam[0])
            assert np.max(np.abs(lam - ref)) <= 1e-13 * ref[0]
            assert np.allclose(W.T @ W, np.eye(k), atol=1e-12)
            back = W @ C if side == 0 else C @ W.T
            assert np.linalg.norm(back - A) <= 1e-12 * np.linalg.norm(A)


@pytest.mark.parametrize("model,L,M", [("HeisenbergU1", 20, 400), ("HeisenbergSU2", 24, 300)])
def test_dmrg_density_matrix_split_matches_svd_split(model, L, M):
    """Sweeps whose split factors on the device through the reduced density
    matrix reach the energy of the same sweeps with the host SVD."""
    be_dev = dmrg.GpuBackend(device_split_min=64)
    r_dev = dmrg.run_dmrg(model, L, M, 2, backend=be_dev, params={"J": 1.0})
    be_host = dmrg.GpuBackend(device_split_min=0)
    r_host = dmrg.run_dmrg(model, L, M, 2, backend=be_host, params={"J": 1.0})
    print(model, "device", r_dev["energy"], "host", r_host["energy"], "device splits", be_dev.device_splits,
          flush=True)
    assert be_dev.device_splits > 0 and be_host.device_splits == 0
    assert abs(r_dev["energy"] - r_host["energy"]) <= 1e-9 * abs(r_host["energy"])
