"""The C60 coupling graph, its Cuthill-McKee chain order and the
general-coupling SU(2) Heisenberg MPO (HeisenbergSU2 with Jfull,
models/HeisenbergSU2.h:223-229; graph from hopping_fullerene,
models/ParamCollection.h:1540-2010).

The graph is pinned by what makes it the truncated icosahedron (12
pentagons, 20 hexagons, 3-regular, the Hueckel spectrum with its HOMO at the
golden ratio - 1); the MPO by exact diagonalisation, and its bond dimensions
by the rank of the couplings across each bond."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import dmrg_oracle  # noqa: E402
import ed  # noqa: E402
from vmps_amd import dmrg, lattice, models  # noqa: E402


def test_c60_is_the_truncated_icosahedron():
    edges = lattice.fullerene_c60_edges()
    assert len(edges) == 90 and len(set(edges)) == 90
    A = lattice.adjacency(edges, 60)
    assert np.all(A.sum(axis=1) == 3)
    assert lattice.polygon_counts(A) == {5: 12, 6: 20}
    ev = np.sort(np.linalg.eigvalsh(A))[::-1]
    vals, counts = np.unique(np.round(ev, 6), return_counts=True)
    spec = dict(zip(vals, counts))
    # Hueckel levels of C60: 3 (a_g), HOMO (sqrt 5 - 1)/2 five-fold (h_u),
    # LUMO three-fold (t_1u) at -0.138564, eigenvalue 1 nine-fold
    assert spec[3.0] == 1
    assert spec[round((np.sqrt(5) - 1) / 2, 6)] == 5
    assert spec[-0.138564] == 3
    assert spec[1.0] == 9
    assert abs(ev.sum()) < 1e-9 and abs((ev ** 2).sum() - 180) < 1e-9  # trace A = 0, trace A^2 = 2 |E|
    # every degeneracy is an icosahedral irrep dimension
    assert set(counts) <= {1, 3, 4, 5, 9}


def test_cuthill_mckee_order():
    A = lattice.adjacency(lattice.fullerene_c60_edges(), 60)
    p = lattice.cuthill_mckee(A)
    assert sorted(p.tolist()) == list(range(60))
    B = lattice.reorder(A, p)
    assert np.allclose(np.sort(np.linalg.eigvalsh(B)), np.sort(np.linalg.eigvalsh(A)), atol=1e-12)
    assert lattice.bandwidth(A) == 15 and lattice.bandwidth(B) == 10
    # breadth-first levels: every vertex's neighbours are at most one level apart
    level = {int(p[0]): 0}
    for v in p[1:]:
        level[int(v)] = min(level[int(u)] for u in np.flatnonzero(A[v]) if int(u) in level) + 1
    for i, j in lattice.fullerene_c60_edges():
        assert abs(level[i] - level[j]) <= 1


def test_c60_mpo_bond_dimensions():
    """D_W = 2 + rank of the couplings across the bond, at most 12 multiplets
    on C60 in Cuthill-McKee order."""
    J = lattice.c60_couplings()
    sites = models.chain_mpo("HeisenbergSU2Full", 60, J=J)
    ranks = lattice.cut_ranks(J)
    assert max(ranks) == 10
    for s, W in enumerate(sites):
        assert W.dw_out == (1 if s == 59 else ranks[s + 1] + 2)
        assert W.dw_in == (1 if s == 0 else ranks[s] + 2)


def test_coupling_channels_reproduce_couplings():
    """The channel matrices compose back to J: J_ij = Q_{i+1}[i] T_{i+1} .. T_{j-1} close_j."""
    rng = np.random.default_rng(5)
    L = 9
    J = np.triu(rng.uniform(-1, 1, (L, L)), 1)
    J[np.abs(J) < 0.3] = 0.0
    J = J + J.T
    ch = models.coupling_channels(J)
    for i in range(L):
        for j in range(i + 1, L):
            x = ch[i][0]
            for s in range(i + 1, j):
                x = x @ ch[s][1]
            assert abs(x @ ch[j][2] - J[i, j]) < 1e-12


@pytest.mark.parametrize("S_tot", [0, 1])
def test_full_coupling_mpo_matches_ed(S_tot):
    """DMRG with the general-coupling MPO (all-to-all random J, channel rank
    up to 4) equals the lowest multiplet of exact diagonalisation."""
    rng = np.random.default_rng(3)
    L = 8
    J = np.triu(rng.uniform(-0.6, 1.0, (L, L)), 1)
    J[np.abs(J) < 0.15] = 0.0
    J = J + J.T
    e_ed = ed.heisenberg_multiplet_spectrum(L, S_tot, Jmat=J)[0]
    r = dmrg.run_dmrg("HeisenbergSU2Full", L, 16, 3, Q=(2 * S_tot + 1,), backend=dmrg_oracle.OracleBackend(),
                      params=dict(J=J))
    assert abs(r["energy"] - e_ed) <= 1e-9 * abs(e_ed)


def test_c60_mid_chain_steps_lower_the_energy():
    """Mid-chain two-site steps on C60 from a random state prepared by
    center_chain (left part left-normalized, right part right-normalized):
    from a random start the step energies fall along the sweep, and stay
    above the bound of -3/4 J per bond."""
    J = lattice.c60_couplings()
    sites = models.chain_mpo("HeisenbergSU2Full", 60, J=J)
    be = dmrg_oracle.OracleBackend()
    ch = dmrg.center_chain(sites, (1,), 12, 29, be, np.random.default_rng(1))
    Es = [dmrg.dmrg_step(ch, be, l, "right")[0] for l in range(29, 33)]
    assert all(b < a + 1e-12 for a, b in zip(Es, Es[1:]))
    assert Es[-1] > -0.75 * 90  # above the bond-wise bound: -3/4 J per bond
