"""The header-only OxV adapter (include/vmps_hxv/pivot_adapter.hpp) driven by
reference-shaped objects: a test program instantiates the reference's generic
HxV -> OxV call pattern on stand-in types with the reference's member names
and checks that overload resolution lands in the adapter.  On the CPU the
device step is absent, so the run must stop in the adapter at plan creation
(after every host-side check passed); on the GPU its Vout must equal the
oracle's."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import synth
from vmps_amd.pivot import get_blockstructure

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pivots  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "adapter", "_build", "adapter_check")


def _q(q):
    return " ".join(str(int(x)) for x in q)


def _qs(qs):
    return "%d %s" % (len(qs), " ".join(_q(q) for q in qs))


def _env(env):
    out = ["%d" % len(env)]
    for (qin, qout, qmid, blocks) in env:
        out.append("%s %s %s %d" % (_q(qin), _q(qout), _q(qmid), len(blocks)))
        for b in blocks:
            if b is None:
                out.append("0 0")
            else:
                out.append("%d %d %s" % (b.shape[0], b.shape[1], " ".join(repr(float(x)) for x in b.ravel(order="F"))))
    return "\n".join(out)


def _W(W):
    out = ["%d" % len(W.blocks)]
    for (s1, s2, k, qi, qo, rows, cols, r, c, v) in W.blocks:
        order = sorted(range(len(v)), key=lambda i: (c[i], r[i]))
        nz = " ".join("%d %d %r" % (r[i], c[i], float(v[i])) for i in order)
        out.append("%d %d %d %s %s %d %d %d %s" % (s1, s2, k, _q(qi), _q(qo), rows, cols, len(v), nz))
    return "\n".join(out)


def write_pivot(pd, vin, path, bs=None):
    lines = ["%d %d %s" % (pd.nsite, pd.sym.nq, " ".join(str(k) for k in pd.sym.kinds))]
    terms = [(pd.W12, pd.W34, pd.L, pd.R)] + [(t.W12, t.W34, t.L, t.R) for t in pd.extra]
    if pd.nsite == 2:
        lines += [_qs(pd.W12.qloc), _qs(pd.W34.qloc), "%d" % len(terms)]
        for (W12, W34, L, R) in terms:
            lines += [_qs(W12.qop), _qs(W34.qop), _env(L), _env(R), _W(W12), _W(W34)]
    elif pd.nsite == 1:
        lines += [_qs(pd.W12.qloc), "%d" % len(terms)]
        for (W12, _, L, R) in terms:
            lines += [_qs(W12.qop), _env(L), _env(R), _W(W12)]
    else:
        lines += [_env(pd.L), _env(pd.R)]
    lines.append("%d" % len(pd.V))
    for blocks in pd.V:
        lines.append("%d %s" % (len(blocks), " ".join("%s %s %d %d" % (_q(a), _q(b), r, c) for (a, b, r, c) in blocks)))
    lines.append(" ".join(repr(float(x)) for x in vin))
    if pd.nsite > 0:
        if bs is None:
            lines.append("0")
        else:
            qlhs, begin, qrhs, fac = bs
            lines.append("%d" % len(qlhs))
            for i in range(len(qlhs)):
                ent = ["%d %d %d" % (qlhs[i][0], qlhs[i][1], begin[i + 1] - begin[i])]
                for j in range(begin[i], begin[i + 1]):
                    ent.append("%s %r" % (" ".join(str(int(x)) for x in qrhs[j]), float(fac[j])))
                lines.append(" ".join(ent))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _oracle_bs(pd):
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    bs = get_blockstructure(lib, h, prefix="oracle_")
    lib.oracle_pivot2_destroy(h)
    return bs


def _cases():
    return [
        ("heis2", False, True), ("hub2", False, True), ("hub2", True, False),
        ("heis1", False, True), ("hub1", True, False), ("hub0", False, False),
    ]


def _prepare(tmp_path, kind, terms, pass_bs):
    pd = pivots.ALL_KINDS[kind]()
    if terms:
        pivots.with_terms(pd, 2)
    v = synth.random_vector(pd, 21)
    bs = _oracle_bs(pd) if pass_bs else None
    src = tmp_path / ("%s.txt" % kind)
    write_pivot(pd, v, str(src), bs)
    return pd, v, src, tmp_path / ("%s.out" % kind)


def _exe():
    if not os.path.exists(EXE):
        pytest.fail("adapter test program not built (run __graft_entry__.build() or make -C tests/adapter)")
    return EXE


@pytest.mark.parametrize("kind,terms,pass_bs", _cases())
def test_adapter_dispatch_cpu(tmp_path, kind, terms, pass_bs):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("device present: covered by the GPU test")
    except ImportError:
        pass
    pd, v, src, dst = _prepare(tmp_path, kind, terms, pass_bs)
    r = subprocess.run([_exe(), str(src), str(dst)], capture_output=True, text=True, timeout=120)
    # 3 = the adapter ran and the library refused at the device step; 2 would
    # mean the generic template won overload resolution
    assert r.returncode == 3, (r.returncode, r.stderr)
    assert r.stderr.startswith("vmps_hxv adapter: plan_create:"), r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("kind,terms,pass_bs", _cases())
def test_adapter_hxv_matches_oracle(tmp_path, kind, terms, pass_bs):
    pd, v, src, dst = _prepare(tmp_path, kind, terms, pass_bs)
    print("case", kind, "terms", terms, "bs", pass_bs, flush=True)
    r = subprocess.run([_exe(), str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    out = np.loadtxt(str(dst))
    ref = pivots.oracle_hxv_terms(oracle.load(), pd, v)
    assert np.linalg.norm(out - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,pass_bs", [("hub2", True), ("heis2", False)])
def test_adapter_inplace_env_change_gets_new_plan(tmp_path, kind, pass_bs):
    """An environment value changed in place (an interior element, so address,
    shape and first/last element stay the same) must not be served from the
    cached plan: the next HxV equals a freshly planned one and differs from
    the one before the change."""
    pd, v, src, dst = _prepare(tmp_path, kind, False, pass_bs)
    print("case inplace", kind, flush=True)
    env = dict(os.environ, VMPS_CHECK_INPLACE="1")
    r = subprocess.run([_exe(), str(src), str(dst)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, (r.returncode, r.stderr)


def test_adapter_complex_calls_keep_generic_templates():
    """Complex seven-argument contract_L / contract_R calls (the TDVP
    instantiation, solvers/TDVPPropagator.h:876-897) compile and bind the
    reference's generic templates, not the FP64 device overloads."""
    r = subprocess.run([_exe(), "--complex-dispatch"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stderr)


# ---- contract_L / contract_R through the seven-argument overloads ----------
def _A(A):
    out = ["%d" % len(A)]
    for blocks in A:
        out.append("%d" % len(blocks))
        for (qin, qout, b) in blocks:
            out.append("%s %s %d %d %s" % (_q(qin), _q(qout), b.shape[0], b.shape[1],
                                          " ".join(repr(float(x)) for x in b.ravel(order="F"))))
    return "\n".join(out)


def write_env_input(which, site, Eold, Abra, Aket, path):
    lines = ["%d %d %s" % (10 + which, site.sym.nq, " ".join(str(k) for k in site.sym.kinds)),
             _qs(site.qloc), _qs(site.qop), _env(Eold), _A(Abra), _A(Aket), _W(site)]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_env_output(path, nq):
    tok = open(path).read().split()
    pos = 0

    def nxt():
        nonlocal pos
        pos += 1
        return tok[pos - 1]
    out = []
    for _ in range(int(nxt())):
        qs = [tuple(int(nxt()) for _ in range(nq)) for _ in range(3)]
        blocks = []
        for _ in range(int(nxt())):
            r, c = int(nxt()), int(nxt())
            blocks.append(np.array([float(nxt()) for _ in range(r * c)]).reshape((r, c), order="F") if r * c else None)
        out.append((qs[0], qs[1], qs[2], blocks))
    return out


def _env_case(tmp_path, which, model):
    from vmps_amd import models
    import envs
    rng = np.random.default_rng(31 + which)
    if model == "heis":
        site = models.heisenberg_u1_mpo()
        left, right = synth.heisenberg_bond(5, 12), synth.heisenberg_bond(6, 12)
    else:
        site = models.hubbard_u1xu1_mpo()
        left = synth.hubbard_bond(3, 16, alpha=1.0, lam=1.2)
        right = synth.hubbard_bond(4, 16, alpha=1.0, lam=1.2)
    Abra = envs.random_site_tensor(site, left, right, rng)
    Aket = envs.random_site_tensor(site, left, right, rng)
    Eold = envs.random_env(site, left if which == 0 else right, rng, "L" if which == 0 else "R")
    src = tmp_path / "env.txt"
    write_env_input(which, site, Eold, Abra, Aket, str(src))
    return site, left, right, Abra, Aket, Eold, src, tmp_path / "env.out"


@pytest.mark.parametrize("which", [0, 1])
def test_adapter_contract_dispatch_cpu(tmp_path, which):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("device present: covered by the GPU test")
    except ImportError:
        pass
    *_, src, dst = _env_case(tmp_path, which, "hub")
    r = subprocess.run([_exe(), str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, (r.returncode, r.stderr)
    assert r.stderr.startswith("vmps_hxv adapter: env_contract:"), r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("model", ["heis", "hub"])
def test_adapter_contract_matches_dense(tmp_path, which, model):
    import envs
    site, left, right, Abra, Aket, Eold, src, dst = _env_case(tmp_path, which, model)
    print("case contract", "LR"[which], model, flush=True)
    r = subprocess.run([_exe(), str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    got = read_env_output(str(dst), site.sym.nq)
    d = len(site.qloc)
    W = site.dense()
    DAb, DAk = envs.dense_A(Abra, left, right, d), envs.dense_A(Aket, left, right, d)
    if which == 0:
        DL = envs.dense_env(Eold, left, site.aux_in)
        ref = np.einsum("sxp,axy,absq,qyr->bpr", DAb, DL, W, DAk)
        out = envs.dense_env(got, right, site.aux_out)
    else:
        DR = envs.dense_env(Eold, right, site.aux_out)
        ref = np.einsum("qxy,byw,abpq,pzw->axz", DAk, DR, W, DAb)
        out = envs.dense_env(got, left, site.aux_in)
    assert np.linalg.norm(out - ref) <= 1e-12 * np.linalg.norm(ref)
