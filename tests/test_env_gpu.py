"""Environment updates contract_L / contract_R on the device against a dense
brute-force contraction (abelian symmetries: all CGC = 1)."""
import numpy as np
import pytest

from vmps_amd import env, models, synth

import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from envs import dense_A, dense_env, random_env, random_site_tensor  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", ["heis", "hub"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("bf16x3", 5e-5)])
def test_contract_L_and_R_vs_dense(model, precision, tol):
    rng = np.random.default_rng(4)
    if model == "heis":
        site = models.heisenberg_u1_mpo()
        left, right = synth.heisenberg_bond(5, 14), synth.heisenberg_bond(6, 14)
    else:
        site = models.hubbard_u1xu1_mpo()
        left = synth.hubbard_bond(3, 20, alpha=1.0, lam=1.2)
        right = synth.hubbard_bond(4, 20, alpha=1.0, lam=1.2)
    d = len(site.qloc)
    Abra = random_site_tensor(site, left, right, rng)
    Aket = random_site_tensor(site, left, right, rng)
    W = site.dense()  # [a, b, s1, s2]
    DAb, DAk = dense_A(Abra, left, right, d), dense_A(Aket, left, right, d)

    Lold = random_env(site, left, rng, "L")
    Lnew = env.contract_L(site, Lold, Abra, Aket, precision=precision)
    DL = dense_env(Lold, left, site.aux_in)
    ref = np.einsum("sxp,axy,absq,qyr->bpr", DAb, DL, W, DAk)  # Abra^T L_a Aket, weighted by W
    got = dense_env(Lnew, right, site.aux_out)
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)

    Rold = random_env(site, right, rng, "R")
    Rnew = env.contract_R(site, Rold, Abra, Aket, precision=precision)
    DR = dense_env(Rold, right, site.aux_out)
    ref = np.einsum("qxy,byw,abpq,pzw->axz", DAk, DR, W, DAb)  # Aket R_b Abra^T
    got = dense_env(Rnew, left, site.aux_in)
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)
