"""GPU FP64 DF-JK through the C ABI vs the CPU oracle (restated MemDFJK)."""
import numpy as np
import pytest

from conftest import announce
from helpers import gpu_jk, maxabs, random_orbitals, water_system

pytestmark = pytest.mark.gpu

TOL = 1e-11  # max-abs, FP64 J/K against the oracle


@pytest.fixture(scope="module")
def h2o4(gpu_available):
    from oracle.dfhelper import MemDFJKOracle
    mol, prim, aux = water_system((2, 2, 1), seed=0)
    announce("(H2O)_4 cc-pVDZ/cc-pVDZ-jkfit: oracle + GPU initialize")
    ora = MemDFJKOracle(prim, aux)
    jk = gpu_jk(prim, aux)
    return mol, prim, aux, ora, jk


def test_symmetric_jk_h2o4(h2o4):
    mol, prim, aux, ora, jk = h2o4
    nbf = prim.nbf()
    C = random_orbitals(nbf, 20, seed=1)
    announce("symmetric J/K nocc=20")
    jk.C_clear()
    jk.C_left_add(C)
    jk.compute()
    D, J, K = ora.compute([C])
    assert maxabs(jk.D()[0], D[0]) < 1e-13
    assert maxabs(jk.J()[0], J[0]) < TOL
    assert maxabs(jk.K()[0], K[0]) < TOL


def test_nonsymmetric_multi_pair_h2o4(h2o4):
    mol, prim, aux, ora, jk = h2o4
    nbf = prim.nbf()
    pairs = [(random_orbitals(nbf, w, seed=10 + w), random_orbitals(nbf, w, seed=100 + w)) for w in (16, 23, 30)]
    announce("non-symmetric J/K, 3 pairs of widths 16/23/30")
    jk.C_clear()
    for L, R in pairs:
        jk.C_left_add(L)
        jk.C_right_add(R)
    jk.compute()
    D, J, K = ora.compute([p[0] for p in pairs], [p[1] for p in pairs])
    for i in range(3):
        assert maxabs(jk.D()[i], D[i]) < 1e-13
        assert maxabs(jk.J()[i], J[i]) < TOL
        assert maxabs(jk.K()[i], K[i]) < TOL
