"""Writes wigner.json: exact 6j and 9j symbols from sympy.physics.wigner for
random admissible arguments (values in units of 1/2, stored as 2j), used to
pin the Racah evaluation of the oracle and, through the block-structure
factors, the engine's.  Run: python tests/golden/make_wigner.py"""
import json
import os
import random

from sympy import Rational, nsimplify
from sympy.physics.wigner import wigner_6j, wigner_9j

HERE = os.path.dirname(os.path.abspath(__file__))


def tri(a, b, c):
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


def admissible6(j):
    a, b, c, d, e, f = j
    return tri(a, b, c) and tri(a, e, f) and tri(d, b, f) and tri(d, e, c)


def admissible9(j):
    a, b, c, d, e, f, g, h, i = j
    return (tri(a, b, c) and tri(d, e, f) and tri(g, h, i) and tri(a, d, g) and tri(b, e, h) and tri(c, f, i))


def main(seed=20261202):
    rng = random.Random(seed)
    out = {"6j": [], "9j": []}
    # 6j: small cases exhaustively-ish plus random large ones (2j up to 60)
    while len(out["6j"]) < 160:
        top = 8 if len(out["6j"]) < 80 else 60
        j = [rng.randint(0, top) for _ in range(6)]
        if not admissible6(j):
            continue
        v = wigner_6j(*[Rational(x, 2) for x in j])
        if v == 0 and rng.random() < 0.7:
            continue
        out["6j"].append([j, str(v), str(v.evalf(30))])
    while len(out["9j"]) < 80:
        top = 6 if len(out["9j"]) < 40 else 24
        j = [rng.randint(0, top) for _ in range(9)]
        if not admissible9(j):
            continue
        v = wigner_9j(*[Rational(x, 2) for x in j], prec=None)
        v = nsimplify(v)
        if v == 0 and rng.random() < 0.7:
            continue
        out["9j"].append([j, str(v), str(v.evalf(30))])
    with open(os.path.join(HERE, "wigner.json"), "w") as f:
        f.write("{\n")
        for n, key in enumerate(("6j", "9j")):
            f.write('"%s": [\n' % key)
            f.write(",\n".join(json.dumps(e) for e in out[key]))
            f.write("\n]%s\n" % ("," if n == 0 else ""))
        f.write("}\n")


if __name__ == "__main__":
    main()
