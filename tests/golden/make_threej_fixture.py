"""Writes threej_sympy.json: exact Wigner 3j values (J 1 J'; k, k'-k, -k')
for every entry of the 3j table at J <= 20 (the entries
precompute_threej fills, trove_functions.cpp:909-919), from
sympy.physics.wigner.wigner_3j, rounded to double.

Run: python tests/golden/make_threej_fixture.py
"""
import json
import os

from sympy.physics.wigner import wigner_3j

JMAX = 20
rows = []
for jI in range(JMAX + 1):
    for jF in range(max(jI - 1, 0), min(jI + 1, JMAX) + 1):
        for kI in range(jI + 1):
            for kF in range(max(kI - 1, 0), min(kI + 1, jF) + 1):
                v = wigner_3j(jI, 1, jF, kI, kF - kI, -kF)
                rows.append([jI, jF, kI, kF, float(v.evalf(30))])
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "threej_sympy.json")
with open(out, "w") as f:
    json.dump({"jmax": JMAX, "source": "sympy.physics.wigner.wigner_3j",
               "columns": ["jI", "jF", "kI", "kF", "value"], "rows": rows}, f)
print(len(rows), "entries ->", out)
