"""Two-site DMRG with the device backend (Lanczos + HxV + contract_L/R on the
GPU): exact diagonalisation, the CPU oracle backend, and mixed precision."""
import sys

import pytest

from vmps_amd import dmrg

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import dmrg_oracle  # noqa: E402
import ed  # noqa: E402

pytestmark = pytest.mark.gpu


def test_gpu_dmrg_matches_ed():
    e_ed = ed.heisenberg_ground(10)
    r = dmrg.run_dmrg("HeisenbergU1", 10, 32, 2, backend=dmrg.GpuBackend())
    assert abs(r["energy"] - e_ed) <= 1e-10 * abs(e_ed)
    e_ed = ed.hubbard_ground(6, U=8.0)
    r = dmrg.run_dmrg("HubbardU1xU1", 6, 64, 2, backend=dmrg.GpuBackend(), params=dict(t=1.0, U=8.0))
    assert abs(r["energy"] - e_ed) <= 1e-10 * abs(e_ed)


def test_gpu_dmrg_matches_oracle_dmrg_L32_M64():
    """HeisenbergU1 L=32, M=64, 4 sweeps from the same random MPS: device and
    CPU-oracle backends give the same energy."""
    r_gpu = dmrg.run_dmrg("HeisenbergU1", 32, 64, 4, backend=dmrg.GpuBackend(tol_eigval=1e-13, tol_state=1e-8))
    r_cpu = dmrg.run_dmrg("HeisenbergU1", 32, 64, 4, backend=dmrg_oracle.OracleBackend())
    print(r_gpu["energy"], r_cpu["energy"], r_gpu["sweep_seconds"], r_cpu["sweep_seconds"])
    assert abs(r_gpu["energy"] - r_cpu["energy"]) <= 1e-9 * abs(r_cpu["energy"])


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_precision_dmrg_energy(prec):
    """Mixed-precision Lanczos with FP64 refinement: |dE| <= 1e-9 vs the FP64 run."""
    r64 = dmrg.run_dmrg("HubbardU1xU1", 12, 128, 3, backend=dmrg.GpuBackend(tol_eigval=1e-12, tol_state=1e-7),
                        params=dict(t=1.0, U=8.0))
    rmx = dmrg.run_dmrg("HubbardU1xU1", 12, 128, 3,
                        backend=dmrg.GpuBackend(precision=prec, tol_eigval=1e-12, tol_state=1e-7, mixed_min_flops=0),
                        params=dict(t=1.0, U=8.0))
    print(prec, r64["energy"], rmx["energy"], rmx["energy"] - r64["energy"])
    assert abs(rmx["energy"] - r64["energy"]) <= 1e-9


def test_gpu_dmrg_su2_matches_ed_and_u1():
    """SU(2)-reduced device DMRG: HeisenbergSU2 L=10 == exact; L=16 with the
    untruncated M = 70 multiplets == U(1) DMRG with the untruncated M = 256
    states (U(1)/SU(2) equality, the reference's own check,
    test_Heisenberg.cc:393)."""
    e_ed = ed.heisenberg_ground(10)
    r = dmrg.run_dmrg("HeisenbergSU2", 10, 16, 2, backend=dmrg.GpuBackend())
    assert abs(r["energy"] - e_ed) <= 1e-10 * abs(e_ed)
    r_su2 = dmrg.run_dmrg("HeisenbergSU2", 16, 70, 3, backend=dmrg.GpuBackend(tol_eigval=1e-13, tol_state=1e-8))
    r_u1 = dmrg.run_dmrg("HeisenbergU1", 16, 256, 3, backend=dmrg.GpuBackend(tol_eigval=1e-13, tol_state=1e-8))
    print("SU2", r_su2["energy"], "U1", r_u1["energy"], "trunc", max(r_su2["truncation"][-15:]),
          max(r_u1["truncation"][-15:]))
    assert abs(r_su2["energy"] - r_u1["energy"]) <= 1e-10 * abs(r_u1["energy"])
    assert abs(r_su2["energy"] - ed.heisenberg_ground(16)) <= 1e-10 * abs(r_u1["energy"])


def test_gpu_dmrg_hubbard_su2xu1_matches_ed():
    e_ed = ed.hubbard_ground(6, U=8.0)
    r = dmrg.run_dmrg("HubbardSU2xU1", 6, 40, 2, backend=dmrg.GpuBackend(), params=dict(t=1.0, U=8.0))
    assert abs(r["energy"] - e_ed) <= 1e-10 * abs(e_ed)


def test_device_environments_give_identical_sweeps():
    """Environments kept on the device (the default) and environments copied
    back to the host after every contraction give the same sweep, bit for bit."""
    kw = dict(tol_eigval=1e-12, tol_state=1e-7)
    print("case HubbardU1xU1 L=8 M=64, device vs host environments", flush=True)
    a = dmrg.run_dmrg("HubbardU1xU1", 8, 64, 2, backend=dmrg.GpuBackend(device_envs=True, **kw),
                      params=dict(t=1.0, U=8.0))
    b = dmrg.run_dmrg("HubbardU1xU1", 8, 64, 2, backend=dmrg.GpuBackend(device_envs=False, **kw),
                      params=dict(t=1.0, U=8.0))
    assert a["energies"] == b["energies"]
    assert a["truncation"] == b["truncation"]
