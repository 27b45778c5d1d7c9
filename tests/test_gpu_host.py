"""The kept front ends on the GPU: the pybind11 CyRSoXS module (notebook-style
flow) and the CyRSoXS executable, each checked against the FP64 oracle."""
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

from cyrsoxs_amd import synthetic
from oracle import oracle as O
from tests.helpers import pattern_rel_l2

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "cyrsoxs_amd")
DATA = os.path.join(REPO, "tests", "golden", "reference_data")
if PKG not in sys.path:
    sys.path.insert(0, PKG)
cy = pytest.importorskip("CyRSoXS")


@pytest.mark.parametrize("precision,tol", [("FP32", 1e-4), ("FP64", 1e-10)])
@pytest.mark.parametrize("pipeline", ["Auto", "MBasis", "EBasis", "Plain"])
def test_notebook_flow_matches_oracle(precision, tol, pipeline):
    n = 32
    vox = synthetic.sphere_morphology(n, 10.0)  # [2][V][4] (S, theta, psi, Vfrac)
    energies = [284.0, 285.0]
    mats = synthetic.sphere_constants(energies)
    inp = cy.InputData(NumMaterial=2)
    inp.setEnergies(energies)
    inp.setERotationAngle(StartAngle=0.0, EndAngle=150.0, IncrementAngle=30.0)
    inp.setMorphologyType(cy.MorphologyType.EulerAngles)
    inp.setPhysSize(5.0)
    inp.setDimensions([n, n, n], cy.MorphologyOrder.ZYX)
    inp.setCaseType(cy.CaseType.Default)
    inp.interpolationType = cy.InterpolationType.NearestNeighour
    inp.precision = getattr(cy.Precision, precision)
    inp.pipeline = getattr(cy.Pipeline, pipeline)
    assert inp.validate()
    ri = cy.RefractiveIndex(inp)
    for e, E in enumerate(energies):
        m = mats[e]
        ri.addData([[1 - m[i, 0], m[i, 1], 1 - m[i, 2], m[i, 3]] for i in range(2)], E)
    vd = cy.VoxelData(InputData=inp)
    for i in range(2):
        f = lambda c: vox[i, :, c].reshape(n, n, n)
        vd.addVoxelData(f(0), f(1), f(2), f(3), i + 1)
    sp = cy.ScatteringPattern(inp)
    cy.launch(inp, ri, vd, sp, WriteMetaData=False)
    got = np.stack([sp.dataToNumpy(E, 0) for E in energies])[:, None]
    ref = O.simulate(O.Sim((n, n, n), 2, 5.0, energies, (0.0, 30.0, 150.0), interpolation=0), vox, mats)
    err, same = pattern_rel_l2(got, ref)
    assert same and err < tol, (err, same)
    # computePolarization (debug entry) against the oracle polarization
    pol = cy.Polarization(inp)
    cy.computePolarization(inp, ri, vd, pol, 285.0, 0.5)
    R = O.rotation_matrix_e((0, 0, 1), 0.5)
    ref_p = O.polarization(O.Sim((n, n, n), 2, 5.0, energies, (0, 1, 0)), vox, mats[1], R)
    for c in range(3):
        g = pol.writeToNumpy(c).reshape(-1)
        r = ref_p[c]
        assert np.linalg.norm(g - r) / np.linalg.norm(r) < (1e-5 if precision == "FP32" else 1e-12)


def _h5dump_projection(path, ny, nx):
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    out = subprocess.run([h5dump, "-d", "/K0/projection", "-y", "-w", "0", "-m", "%.9g", path], capture_output=True,
                         text=True).stdout
    start = out.index("DATA {") + 6
    body = out[start:out.index("}", start)]
    vals = [float(x) for x in re.findall(r"-?nan|[-+]?[0-9.]+(?:[eE][-+]?[0-9]+)?", body)]
    return np.array(vals).reshape(ny, nx)


@pytest.mark.skipif(not cy.hdf5Available(), reason="libhdf5 not available")
def test_cli_on_reference_data(tmp_path):
    """The reference's Data/ inputs (vector morphology, 280 eV, FULL approach, NN)
    through the executable.  Data/ ships Material0.txt and Material1.txt while the
    reader opens Material1..N.txt (include/Input/InputData.h:343), so they are
    staged as Material1.txt and Material2.txt.  (Material1.txt alone for both would
    make the two materials identical and isotropic at 280 eV: a contrast-free
    morphology whose pattern is rounding noise.)"""
    shutil.copy(os.path.join(DATA, "config.txt"), tmp_path)
    shutil.copy(os.path.join(DATA, "Material0.txt"), tmp_path / "Material1.txt")
    shutil.copy(os.path.join(DATA, "Material1.txt"), tmp_path / "Material2.txt")
    h5 = os.path.join(DATA, "edgeSphereZYX.h5")
    exe = os.path.join(PKG, "lib", "CyRSoXS")
    r = subprocess.run([exe, h5, "out", "--precision", "fp64"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "CyRSoXS.log").exists()
    got = _h5dump_projection(str(tmp_path / "out" / "Energy_280.00.h5"), 32, 32)
    cfg = cy.readConfig(str(tmp_path / "config.txt"), 2, str(tmp_path))
    inp = cy.InputData(NumMaterial=2)
    inp.setEnergies([280.0])
    inp.setERotationAngle(0.0, 0.0, 1.0)
    inp.setMorphologyType(cy.MorphologyType.VectorMorphology)
    inp.setPhysSize(5.0)
    inp.setDimensions([16, 32, 32], cy.MorphologyOrder.ZYX)
    inp.setCaseType(cy.CaseType.Default)
    vd = cy.VoxelData(inp)
    vd.readFromH5(h5)
    vox = vd.dataToNumpy()
    m = np.array([[[c[0].real, c[0].imag, c[1].real, c[1].imag] for c in cfg["Materials"]]])
    ref = O.simulate(O.Sim((32, 32, 16), 2, 5.0, [280.0], (0.0, 1.0, 0.0), morphology_type=1, interpolation=0,
                           scatter_approach=1), vox, m)[0, 0]
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(got), fin)
    assert np.linalg.norm(got[fin] - ref[fin]) / np.linalg.norm(ref[fin]) < 1e-8  # h5dump prints 9 digits
