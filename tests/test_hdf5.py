"""HDF5 spin files (helpers/load.h:13-19, :53-57) and the "hdf5" monitor
(monitors/hdf5.cc:27-289) through the host layer.

The files are checked with the HDF5 tools of the image (h5import writes a
spin file independently of this build, h5dump reads what the monitor wrote);
the tests are skipped where the tools or the HDF5 C library are absent.
"""
import os
import re
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from jams_amd import configs
from jams_amd.host import Simulation


def _tool(name):
    p = shutil.which(name) or (f"/opt/conda/bin/{name}" if os.path.exists(f"/opt/conda/bin/{name}") else None)
    if p is None:
        pytest.skip(f"{name} is not available")
    return p


def _h5import_spins(tmp_path, s):
    """An HDF5 file with the N x 3 dataset /spins written by h5import from
    text (lossless %.17g)."""
    txt = tmp_path / "spins.txt"
    np.savetxt(txt, s, fmt="%.17g")
    cfg = tmp_path / "spins.cfg"
    cfg.write_text(f"PATH spins\nINPUT-CLASS TEXTFP\nINPUT-SIZE 64\nRANK 2\n"
                   f"DIMENSION-SIZES {s.shape[0]} 3\n"
                   "OUTPUT-CLASS FP\nOUTPUT-SIZE 64\nOUTPUT-ARCHITECTURE NATIVE\n")
    out = tmp_path / "spins.h5"
    subprocess.run([_tool("h5import"), str(txt), "-c", str(cfg), "-o", str(out)], check=True,
                   capture_output=True)
    return out


def _h5dump_values(path, dataset):
    """The values of a dataset as h5dump prints them at full precision."""
    r = subprocess.run([_tool("h5dump"), "-m", "%.17g", "-d", dataset, str(path)], check=True,
                       capture_output=True, text=True)
    body = r.stdout.split("DATA {", 1)[1]
    body = re.sub(r"\(\d+(,\d+)*\):", " ", body)
    return np.array([float(x) for x in re.findall(r"[-+0-9.eEinfa]+", body.split("}")[0])])


def test_lattice_spins_from_hdf5_file(tmp_path):
    """lattice.spins = "<file>.h5" loads the dataset /spins (written here by
    h5import, not by this build), as the reference's load_array_from_file
    does; a file of the wrong size is refused."""
    rng = np.random.default_rng(5)
    s = rng.normal(size=(4 * 4 * 4, 3))
    s /= np.linalg.norm(s, axis=1)[:, None]
    f = _h5import_spins(tmp_path, s)
    text = configs.sc_config(L=(4, 4, 4), steps=10).replace(
        "periodic = [true, true, true];", f'periodic = [true, true, true]; spins = "{f}";')
    sim = Simulation(text, device=-1, output_dir=str(tmp_path), name="h5in")
    assert np.array_equal(sim.sites()[3], s)
    sim.close()
    bad = configs.sc_config(L=(4, 4, 5), steps=10).replace(
        "periodic = [true, true, true];", f'periodic = [true, true, true]; spins = "{f}";')
    with pytest.raises(RuntimeError, match="expected size"):
        Simulation(bad, device=-1, output_dir=str(tmp_path), name="h5bad")


@pytest.mark.gpu
def test_hdf5_monitor_files_and_reload(tmp_path):
    """An sc 6^3 run (a = 0.3 nm, random start, 30 steps) with the hdf5
    monitor every 10 steps: <sim>_lattice.h5 holds the material ids, the
    moments (in meV/T, as the lattice holds them) and the positions in nm
    (attribute units = "nm");
    <sim>_0000000.h5 .. _0000003.h5 and <sim>_final.h5 hold /spins, the last
    ones equal to the engine's spins bit for bit (read back with h5dump);
    the XDMF time series lists each snapshot with its time; and a new
    simulation started from lattice.spins = <sim>_final.h5 has those spins."""
    mon = 'monitors = ( { module = "hdf5"; output_steps = 10; compressed = true; } );'
    text = configs.sc_config(L=(6, 6, 6), steps=30, monitors=mon)
    sim = Simulation(text, device=0, output_dir=str(tmp_path), name="snap")
    sim.run()
    final = sim.engine().get_spins()
    mus = sim.sites()[0]
    sim.close()
    n = 6 * 6 * 6
    lat = tmp_path / "snap_lattice.h5"
    pos = _h5dump_values(lat, "/positions").reshape(n, 3)
    ijk = np.array([(i, j, k) for i in range(6) for j in range(6) for k in range(6)], float)
    assert np.allclose(pos, ijk * 0.3, rtol=0, atol=1e-12)
    assert np.array_equal(_h5dump_values(lat, "/types"), np.zeros(n))
    # the moments as the lattice holds them (2 muB in meV/T, as globals::mus)
    assert np.array_equal(_h5dump_values(lat, "/moments"), mus)
    assert np.allclose(mus, 2.0 * 0.0578838181, rtol=1e-9)
    attr = subprocess.run([_tool("h5dump"), "-a", "/positions/units", str(lat)], check=True,
                          capture_output=True, text=True).stdout
    assert '"nm"' in attr
    for k in range(4):
        assert (tmp_path / f"snap_{k:07d}.h5").exists(), k
    for name in ("snap_0000003.h5", "snap_final.h5"):
        got = _h5dump_values(tmp_path / name, "/spins").reshape(n, 3)
        assert np.array_equal(got, final), name
    root = ET.parse(tmp_path / "snap.xdmf").getroot()
    grids = root.find("Domain").find("Grid").findall("Grid")
    assert len(grids) == 5  # 0, 10, 20, 30 and the final snapshot
    times = [float(g.find("Time").get("Value")) for g in grids]
    assert times[:4] == pytest.approx([0.0, 0.01, 0.02, 0.03], abs=1e-9)
    assert grids[-1].find("Attribute[@Name='spin']/DataItem").text.strip() == "snap_final.h5:/spins"
    again = configs.sc_config(L=(6, 6, 6), steps=30).replace(
        "periodic = [true, true, true];",
        f'periodic = [true, true, true]; spins = "{tmp_path / "snap_final.h5"}";')
    sim2 = Simulation(again, device=-1, output_dir=str(tmp_path), name="again")
    assert np.array_equal(sim2.sites()[3], final)
    sim2.close()
