"""Batched HxV: several independent pivots (different models, symmetries and
sizes) in one plan; each pivot's slice of the result equals the oracle's HxV
of that pivot alone."""
import os
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import synth
from vmps_amd.pivot import HxV, PivotBatch

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pivots  # noqa: E402
from test_su2_cpu import su2_pivot  # noqa: E402

pytestmark = pytest.mark.gpu


def _oracle(pd, v):
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    try:
        return pivots.oracle_hxv(lib, h, v)
    finally:
        lib.oracle_pivot2_destroy(h)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("bf16x3", 5e-5)])
def test_batch_equals_individual_pivots(precision, tol):
    pds = [pivots.small_heisenberg(), pivots.small_hubbard(), su2_pivot(), pivots.small_hubbard(seed=9, M=24),
           synth.make_workload("heisenberg_u1_L32_M64")]
    print("batch of", len(pds), precision, flush=True)
    B = PivotBatch(pds, precision=precision)
    assert B.vsize == sum(pd.vsize for pd in pds)
    v = np.concatenate([synth.random_vector(pd, 3 + i) for i, pd in enumerate(pds)])
    out = HxV(B, v)
    for pd, vi, oi in zip(pds, B.split(v), B.split(out)):
        ref = _oracle(pd, vi)
        assert np.linalg.norm(oi - ref) <= tol * np.linalg.norm(ref)
