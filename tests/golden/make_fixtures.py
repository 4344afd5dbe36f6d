"""Extract the reference test fixtures (data only) into tests/golden/*.json.

Run in the development container, where the reference tree is mounted:
    python tests/golden/make_fixtures.py /root/reference

It copies numbers, not code: the hard-coded integral maps of the
reference's test fixtures, the benzene FCIDUMP example, the 24 Wigner-9j
test values, the golden energies and the MPO shape figures printed by the
benzene two-site example run.  The sources are cited in each record.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def integral_map(path, start_pat):
    text = open(path).read()
    m = re.search(start_pat, text)
    if not m:
        raise RuntimeError(f"integral map not found in {path}")
    i = m.end()
    depth = 1
    j = i
    while depth:
        if text[j] == "{":
            depth += 1
        elif text[j] == "}":
            depth -= 1
        j += 1
    body = text[i:j - 1]
    ents = re.findall(r"\{\s*\{\s*(-?\d+)\s*,\s*(-?\d+)\s*,\s*(-?\d+)\s*,\s*(-?\d+)\s*\}\s*,\s*([-+0-9.eE]+)\s*\}", body)
    return [[[int(a), int(b), int(c), int(d)], float(v)] for a, b, c, d, v in ents]


def main(ref):
    t = os.path.join(ref, "dmrg", "tests")
    out = {}
    out["h2"] = dict(
        source="dmrg/tests/Fixtures/H2Fixture.h:25-36; energy dmrg/tests/test1.cpp:93",
        L=2, nelec=2, spin=0, nup=1, ndown=1,
        integrals=integral_map(os.path.join(t, "Fixtures", "H2Fixture.h"), r"integralH2\s*=\s*RealIntegralMapType\s*\{"),
        energy=-0.980724992658492, rtol=1e-9)
    out["lih"] = dict(
        source="dmrg/tests/Fixtures/LiHFixture.h:28-112; energies dmrg/tests/test2.cpp:57,139",
        L=4, nelec=2, spin=0, nup=1, ndown=1,
        integrals=integral_map(os.path.join(t, "Fixtures", "LiHFixture.h"), r"integralsLiH\s*=\s*RealIntegralMapType\s*\{"),
        energy=-7.90435750473166, energy_excited_su2=-7.7734906999491207,
        energy_excited_2u1=-7.78681053707699, rtol=1e-9)
    out["siteproblem"] = dict(
        source="dmrg/tests/test_mps_mpo_ops/test_siteproblem_fixture.h:21-87",
        L=4, nelec=2, spin=0, nup=1, ndown=1,
        integrals=integral_map(os.path.join(t, "test_mps_mpo_ops", "test_siteproblem_fixture.h"),
                               r"maquis::integral_map<double>\s*\{"))
    out["benzene"] = dict(
        source="dmrg/tests/Fixtures/BenzeneFixture.h:40-; energies dmrg/tests/FEAST/ElectronicFEAST.cpp:192,249-250",
        L=6, nelec=6, spin=0, nup=3, ndown=3,
        integrals=integral_map(os.path.join(t, "Fixtures", "BenzeneFixture.h"), r"integralsReal\s*=\s*RealIntegralMapType\s*\{"),
        energy=-230.7555388354673, energy_excited=-230.4936880170761, rtol=1e-10)
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)

    # benzene FCIDUMP example and the MPO shape of its two-site run
    src = os.path.join(ref, "examples", "TD-DMRG", "FCIDUMP_Benzene_631Gs_CAS6-6")
    with open(src) as f:
        fcidump = f.read()
    with open(os.path.join(HERE, "FCIDUMP_benzene_cas66"), "w") as f:
        f.write(fcidump)
    log = open(os.path.join(ref, "examples", "TD-DMRG", "Benzene_6e6o.TD.TS.out")).read()
    shape = dict(
        source="examples/TD-DMRG/Benzene_6e6o.TD.TS.out:15-28 (symmetry su2u1pg, FCIDUMP_Benzene_631Gs_CAS6-6)",
        n_terms=int(re.search(r"hamiltonian will contain (\d+) terms", log).group(1)),
        bonds=[[int(a), int(b)] for a, b in re.findall(r"MPO Bond \d+: (\d+)/(\d+)", log)],
        ts_mpo=[[int(a), int(b), int(c)] for a, b, c in re.findall(r"TSMPOTensor: (\d+)x(\d+),\s+(\d+) operators", log)])
    log2 = open(os.path.join(ref, "examples", "iTD-DMRG", "H2_2e4o.TI.SS.out")).read()
    shape_h2 = dict(
        source="examples/iTD-DMRG/H2_2e4o.TI.SS.out:15-19 (symmetry 2u1pg, FCIDUMP_H2_TI)",
        n_terms=int(re.search(r"hamiltonian will contain (\d+) terms", log2).group(1)),
        bonds=[[int(a), int(b)] for a, b in re.findall(r"MPO Bond \d+: (\d+)/(\d+)", log2)])
    with open(os.path.join(ref, "examples", "iTD-DMRG", "FCIDUMP_H2_TI")) as f:
        with open(os.path.join(HERE, "FCIDUMP_h2_cas24"), "w") as g:
            g.write(f.read())
    # Wigner 9j table
    wt = open(os.path.join(t, "test_wigner.cpp")).read()
    cases = re.findall(r"\{\s*\{([-0-9,\s]+)\}\s*,\s*([-+0-9.eE]+)\s*\}", wt)
    nine = [[[int(x) for x in a.replace(" ", "").split(",") if x], float(v)] for a, v in cases]
    with open(os.path.join(HERE, "shapes_and_wigner.json"), "w") as f:
        json.dump(dict(benzene_su2_ts=shape, h2_2u1=shape_h2,
                       wigner9j=dict(source="dmrg/tests/test_wigner.cpp:21-45", rtol=1e-8, cases=nine)), f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
