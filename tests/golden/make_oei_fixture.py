"""Extract psi4's stored angular-momentum integrals of H2O / cc-pVDZ into a fixture.

Source: psi4 tests/pytests/oei_reference_data.json (psi4 1.5, used by
tests/pytests/test_mcmurchie_davidson.py:15-84 at atol 1e-14).  Only the data of
the water / cc-pVDZ entry is kept: the geometry (bohr, no reorientation) and
MintsHelper::ao_angular_momentum() as a 3 x 24 x 24 array.

Usage (in a checkout that has the reference):  python tests/golden/make_oei_fixture.py
"""
import json
import os
import sys

SRC = "/root/reference/tests/pytests/oei_reference_data.json"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "psi4_ao_angular_momentum_h2o_ccpvdz.json")


def main():
    d = json.load(open(SRC))
    md = d["data"]["h2o"]
    out = {
        "source": "psi4 tests/pytests/oei_reference_data.json (version %s, %s), entry h2o / cc-pvdz" % (d["version"], d["date"]),
        "geometry_bohr": [["O", 0.0, 0.0, -0.149544924691], ["H", 0.0, -1.336238163149, 1.186693238458],
                          ["H", 0.0, 1.336238163149, 1.186693238458]],
        "psi4string": md["psi4string"],
        "basis": "cc-pvdz",
        "L_shape": md["cc-pvdz"]["L_shape"],
        "L": md["cc-pvdz"]["L"],
    }
    json.dump(out, open(DST, "w"))
    print("wrote", DST)


if __name__ == "__main__":
    sys.exit(main())
