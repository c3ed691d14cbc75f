"""iga_intdensity on the GPU (igap_gpu_intdensity) against the oracle on the
parity cases: energies and stress integrals within 1e-12 relative, the phase
fractions within one Gauss point's weight (an indicator is a step function of
the strain; a rounding-level difference at a point on a phase boundary moves
one weight), the transient volume equal to the patch volume, and the
Neumann work of the traction case."""
import numpy as np
import pytest

from test_gpu_parity import build_case, cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [c[0] for c in cases()])
def test_intdensity_matches_oracle(igap, oracle, case):
    c = [x for x in cases() if x[0] == case][0]
    mesh_o, mesh_g, ph_o, ph_g, bc_o, bc_g, U = build_case(igap, oracle, c)
    ref = oracle.intdensity(mesh_o, ph_o, bc_o, U)
    with igap.Assm(mesh_g, ph_g, bc_g) as A:
        A.set_uhist(U)
        HH = A.intdensity()
    print(f"[{case}] HH = {np.array2string(HH, precision=6)}")
    assert HH.shape == ref.shape
    nc = 7 if HH.size == 13 else 2
    scale = np.max(np.abs(ref[:nc]))
    assert np.max(np.abs(HH[:nc] - ref[:nc])) <= 1e-12 * scale
    if HH.size == 13:
        wmax = oracle.volume(mesh_o) / (64 * int(np.prod(c[1]))) * 4   # generous bound on one Gauss weight
        assert np.max(np.abs(HH[7:] - ref[7:])) <= wmax
    else:
        assert abs(HH[1] - oracle.volume(mesh_o)) < 1e-12


def test_volume_kat_unit_cube(igap, oracle):
    """ex1's KAT: MeshIGABeam(1,1,1, mref = 3, porder 2, periodic in all three
    directions) (example/ex1/main.c:23; mesh/igabeam.c:30 gives 2^mref = 8
    elements per direction) with gradelasttime and the IGAPSolvRand state:
    the second density integrates to the volume 1."""
    m = igap.MeshIGABeam(L=(1, 1, 1), nelem=(8, 8, 8), periodic=(1, 1, 1))
    with igap.Assm(m, igap.PhysGradElastTime(), igap.BCDefault()) as A:
        A.set_uhist(oracle.glibc_rand_state(3 * A.ndof))
        HH = A.intdensity()
    assert abs(HH[1] - 1.0) < 1e-13
