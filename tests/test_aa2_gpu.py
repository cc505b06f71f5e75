"""contract_AA2 on the device (vmps_contract_aa2_device) against the host
restatement two_site_psi, whose split <-> contract round trip the CPU suite
pins, on DMRG states of U(1)xU(1), SU(2) and SU(2)xU(1) chains."""
import numpy as np
import pytest

from vmps_amd import _lib, dmrg, models
from vmps_amd.pivot import DeviceVector, PivotMatrix2, contract_aa2

pytestmark = pytest.mark.gpu

CASES = [
    ("HubbardU1xU1", 12, (12, 0), [32, 64], 5, dict(t=1.0, U=8.0)),
    ("HeisenbergSU2", 16, (1,), [16, 32], 7, {}),
    ("HubbardSU2xU1", 10, (1, 10), [16, 32], 4, dict(t=1.0, U=6.0)),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def state(request):
    model, L, Q, sched, p, params = request.param
    sites = models.chain_mpo(model, L, **params)
    be = dmrg.GpuBackend(tol_eigval=1e-10, tol_state=1e-6)
    chain = dmrg.warm_chain(sites, Q, sched, p, be, np.random.default_rng(11))
    return chain, p


def _device_psi(pd, terms):
    H = PivotMatrix2(pd)
    try:
        dv = DeviceVector(H)
        contract_aa2(H, dv, *terms)
        out = dv.download()
        dv.free()
    finally:
        H.close()
    return out


def test_device_aa2_matches_host(state):
    chain, p = state
    pd = chain.pivot(p)
    ref = dmrg.two_site_psi(pd, chain.A[p], chain.A[p + 1], chain.qloc)
    terms = dmrg.two_site_psi_terms(pd, chain.A[p], chain.A[p + 1], chain.qloc)
    out = _device_psi(pd, terms)
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    print("contract_AA2 vsize", pd.vsize, "terms", len(terms[2]), "relative error %.2e" % err, flush=True)
    assert err <= 1e-14
    # entries no product reaches are exactly zero on both sides
    touched = np.zeros(pd.vsize, dtype=bool)
    for o, a, b in zip(terms[2], terms[3], terms[4]):
        touched[o:o + terms[0][a].shape[0] * terms[1][b].shape[1]] = True
    assert np.all(out[~touched] == 0.0) and np.all(ref[~touched] == 0.0)


def test_device_aa2_linear_in_factors(state):
    """Doubling every factor doubles the vector; dropping all terms gives zero."""
    chain, p = state
    pd = chain.pivot(p)
    b1, b2, off, i1, i2, f = dmrg.two_site_psi_terms(pd, chain.A[p], chain.A[p + 1], chain.qloc)
    one = _device_psi(pd, (b1, b2, off, i1, i2, f))
    two = _device_psi(pd, (b1, b2, off, i1, i2, [2.0 * x for x in f]))
    assert np.array_equal(two, 2.0 * one)
    assert not np.any(_device_psi(pd, (b1, b2, [], [], [], [])))


def test_device_aa2_rejects_bad_terms(state):
    chain, p = state
    pd = chain.pivot(p)
    b1, b2, off, i1, i2, f = dmrg.two_site_psi_terms(pd, chain.A[p], chain.A[p + 1], chain.qloc)
    H = PivotMatrix2(pd)
    dv = DeviceVector(H)
    try:
        with pytest.raises(_lib.VmpsStatusError):  # block outside the vector
            contract_aa2(H, dv, b1, b2, [pd.vsize], i1[:1], i2[:1], f[:1])
        with pytest.raises(_lib.VmpsStatusError):  # index out of range
            contract_aa2(H, dv, b1, b2, off[:1], [len(b1)], i2[:1], f[:1])
        bad = [j for j in range(len(b2)) if b2[j].shape[0] != b1[i1[0]].shape[1]]
        if bad:  # operands that do not chain
            with pytest.raises(_lib.VmpsStatusError):
                contract_aa2(H, dv, b1, b2, off[:1], i1[:1], bad[:1], f[:1])
    finally:
        dv.free()
        H.close()


def test_dmrg_step_device_psi_matches_host_psi(state):
    """One DMRG step through the device start vector and one through the host
    start vector, from copies of the same state: the same energy."""
    chain, p = state
    be = dmrg.GpuBackend(tol_eigval=1e-12, tol_state=1e-8, device_psi_min=0)

    class HostPsi:  # the same backend without the device start-vector path
        def __init__(self, b):
            self.b = b

        def __getattr__(self, k):
            if k == "solve_two_site":
                raise AttributeError(k)
            return getattr(self.b, k)

    e_dev = dmrg.dmrg_step(chain.copy(), be, p, "right")[0]
    e_host = dmrg.dmrg_step(chain.copy(), HostPsi(be), p, "right")[0]
    print("step energy device psi %.14f host psi %.14f" % (e_dev, e_host), flush=True)
    assert abs(e_dev - e_host) <= 1e-10 * abs(e_host)


def test_device_gather_split_matches_host_split(state):
    """The split of the eigenvector left on the device (sector matrices built
    by vmps_block_eigh_split_gather) against the split of its host copy: the
    same kept dimensions and truncation weight, and the same two-site
    wavefunction rebuilt from the new site tensors."""
    chain, p = state
    pd = chain.pivot(p)
    be = dmrg.GpuBackend(tol_eigval=1e-12, tol_state=1e-8, device_split_min=1)
    E, dev, info = be.solve_two_site(pd, chain.A[p], chain.A[p + 1], chain.qloc, seed=1, keep_device=True)
    try:
        v = dev.download()
        M = max(8, sum(chain.bonds[p + 1].values()) // 2)  # truncating
        host = dmrg.split_psi(pd, v, chain.qloc, M, eigh=be.eigh)
        devs = dmrg.split_psi(pd, dev, chain.qloc, M, eigh=be.eigh, eigh_gather=be.eigh_gather)
    finally:
        dev.close()
    assert host[4] == devs[4]
    assert abs(host[2] - devs[2]) <= 1e-12 * max(1.0, abs(host[2])) + 1e-15
    psi_h = dmrg.two_site_psi(pd, host[0], host[1], chain.qloc)
    psi_d = dmrg.two_site_psi(pd, devs[0], devs[1], chain.qloc)
    err = np.linalg.norm(psi_d - psi_h) / np.linalg.norm(psi_h)
    print("gather split: kept", sum(devs[4].values()), "truncation %.3e" % devs[2], "psi difference %.2e" % err, flush=True)
    # the two density matrices differ by rounding (fused multiply-adds on the
    # device), which turns the kept states near the cut by rounding / gap; the
    # rebuilt wavefunctions agree to that level, and each misses the full one
    # by exactly its truncation weight (weighted norm)
    assert err <= 1e-9
    w = np.concatenate([np.full(r * c, pd.sym.coeff_dot(qo)) for (_, _, qo, _, r, c) in pd.block_offsets()])
    nrm = np.sum(w * v * v)
    for psi, res in ((psi_h, host), (psi_d, devs)):
        assert abs(np.sum(w * (psi - v) ** 2) - res[2]) <= 1e-12 * nrm


def test_device_gather_split_rejects_bad_tables(state):
    chain, p = state
    pd = chain.pivot(p)
    be = dmrg.GpuBackend(tol_eigval=1e-10, tol_state=1e-6)
    E, dev, info = be.solve_two_site(pd, chain.A[p], chain.A[p + 1], chain.qloc, seed=1, keep_device=True)
    try:
        with pytest.raises(_lib.VmpsStatusError):  # block outside its sector
            dmrg.block_eigh_gather_device(dev, [(2, 2)], [(0, 1, 0, 2, 2, [(0, 1.0)])], 0)
        with pytest.raises(_lib.VmpsStatusError):  # part outside the vector
            dmrg.block_eigh_gather_device(dev, [(2, 2)], [(0, 0, 0, 2, 2, [(pd.vsize - 1, 1.0)])], 0)
        with pytest.raises(_lib.VmpsStatusError):  # sector index out of range
            dmrg.block_eigh_gather_device(dev, [(2, 2)], [(1, 0, 0, 2, 2, [(0, 1.0)])], 0)
    finally:
        dev.close()
