"""Solution files in the reference's format (IGAPSolvSaveSolution /
IGAPSolvLoadSolution, src/interface.c:271-300; src/ioutil.c:90-159): raw
FP64 in the universal dof order, static level 0, transient levels 1..2, and a
transient load sets level 0 = level 1."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_static_roundtrip(igap, tmp_path):
    m = igap.MeshIGABeam(L=(4, 1, 1), nelem=(6, 3, 4))
    with igap.Assm(m, igap.PhysGradElast(), igap.BCDefault()) as A:
        U = np.random.default_rng(0).standard_normal(A.ndof)
        A.set_uhist(U)
        f = tmp_path / "u.bin"
        A.save_solution(f)
        raw = np.fromfile(f, dtype=np.float64)
        assert raw.size == A.ndof and np.array_equal(raw, U)
        A.set_uhist(np.zeros(A.ndof))
        A.load_solution(f)
        assert np.array_equal(A.get_uhist(), U)


def test_transient_levels(igap, tmp_path):
    m = igap.MeshIGABeam(L=(1, 1, 1), nelem=(5, 5, 5), periodic=(1, 1, 1))
    with igap.Assm(m, igap.PhysGradElastTime(), igap.BCDefault()) as A:
        n = A.ndof
        H = np.random.default_rng(1).standard_normal(3 * n)
        A.set_uhist(H)
        f = tmp_path / "u.bin"
        A.save_solution(f)
        raw = np.fromfile(f, dtype=np.float64)
        assert raw.size == 2 * n and np.array_equal(raw, H[n:])
        A.set_uhist(np.zeros(3 * n))
        A.load_solution(f)
        G = A.get_uhist()
        assert np.array_equal(G[n:], H[n:]) and np.array_equal(G[:n], H[n:2 * n])


def test_short_file_is_an_error(igap, tmp_path):
    from igap4_amd._lib import IgapError
    m = igap.MeshIGABeam(L=(4, 1, 1), nelem=(4, 3, 3))
    with igap.Assm(m, igap.PhysGradElast(), igap.BCDefault()) as A:
        f = tmp_path / "short.bin"
        np.zeros(A.ndof - 1).tofile(f)
        with pytest.raises(IgapError, match="holds"):
            A.load_solution(f)
