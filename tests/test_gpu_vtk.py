"""VTK output as the reference writes it for the immiscible model (VtkMultiWriter with
the default fields of VtkMultiPhaseModule, io/vtkmultiphasemodule.hh:61-72, :311-337):
per call one "<sim>-NNNNN.vtu" with the cell fields, and a "<sim>.pvd" collection of
every file written. Fields are checked cell by cell: saturations and porosity against the
state, the non-wetting pressure and relative permeabilities against the oracle's material
laws at the same saturations."""
import xml.etree.ElementTree as ET

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ["pressure_w", "pressure_n", "density_w", "density_n", "saturation_w",
          "saturation_n", "relativePerm_w", "relativePerm_n", "porosity"]


def _arrays(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {}
    for da in piece.iter("DataArray"):
        out[da.get("Name")] = np.array(da.text.split(), dtype=np.float64)
    return piece, out


@pytest.mark.parametrize("which", ["lens", "het3d"])
def test_vtk_fields_and_collection(tmp_path, which):
    import ewg
    from ewg import problems
    from oracle.pyoracle import OracleContext
    P = (problems.lens_problem(nx=12, ny=8) if which == "lens"
         else problems.heterogeneous_3d(9, 7, 5, seed=3, patch=2))
    g = ewg.GpuContext(P)
    o = OracleContext(P)
    times = []
    for step in range(2):
        assert g.newton(P.dt)[0] == 0
        g.advance_time_level()
        times.append((step + 1) * P.dt)
        g.write_vtk(tmp_path, "case", times[-1])
    files = [tmp_path / f"case-{i:05d}.vtu" for i in range(2)]
    assert all(f.exists() for f in files)
    piece, a = _arrays(files[1])
    assert int(piece.get("NumberOfCells")) == P.N
    npts = (P.nx + 1) * (P.ny + 1) * ((P.nz + 1) if P.dim == 3 else 1)
    assert int(piece.get("NumberOfPoints")) == npts
    assert [n for n in a if n in FIELDS] == FIELDS
    x = g.solution(0).reshape(-1, 2)
    assert np.array_equal(a["saturation_w"], x[:, 1])
    assert np.array_equal(a["saturation_n"], 1.0 - x[:, 1])
    assert np.array_equal(a["pressure_w"], x[:, 0])
    assert np.array_equal(a["porosity"], P.phi)
    assert np.all(a["density_w"] == P.rho[0]) and np.all(a["density_n"] == P.rho[1])
    pe = P.pe if P.pe is not None else np.zeros(P.N)
    mat = np.array([o.material(int(P.region[i]), x[i, 1], pe=float(pe[i])) for i in range(P.N)])
    assert np.abs(a["pressure_n"] - (x[:, 0] + mat[:, 0])).max() <= 1e-12 * np.abs(x[:, 0]).max()
    assert np.abs(a["relativePerm_w"] - mat[:, 2]).max() <= 1e-14
    assert np.abs(a["relativePerm_n"] - mat[:, 4]).max() <= 1e-14
    # geometry: cell 0 spans the first corner's vertices
    corners = 8 if P.dim == 3 else 4
    conn = a["connectivity"].astype(np.int64)
    assert conn.size == corners * P.N and conn[:4].tolist() == [0, 1, P.nx + 2, P.nx + 1]
    assert np.all(a["types"] == (12 if P.dim == 3 else 9))
    coords = a["Coordinates"].reshape(-1, 3)
    assert np.allclose(coords[-1], [P.lx, P.ly, P.lz if P.dim == 3 else 0.0])
    # the collection lists both files with their times
    root = ET.parse(tmp_path / "case.pvd").getroot()
    ds = root.findall("Collection/DataSet")
    assert [d.get("file") for d in ds] == [f.name for f in files]
    assert [float(d.get("timestep")) for d in ds] == times
