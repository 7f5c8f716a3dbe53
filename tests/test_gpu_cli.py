"""GPU: the hard-wall command line end to end (reference
``test/commandline/test_hard_wall.py``): pressure, displacement and automatic
programmes, text maps and NetCDF frames equal to direct solves."""

import numpy as np
import pytest

from contactmechanics_amd import PeriodicFFTElasticHalfSpace, Topography, make_system
from contactmechanics_amd.CommandLineInterface import HardWall
from contactmechanics_amd.IO.NetCDF import NetCDFContainer
from contactmechanics_amd.Tools.Logger import quiet

pytestmark = pytest.mark.gpu


def sphere_file(tmp_path, n=64, R=2.0):
    x = (np.arange(n) + 0.5) / n - 0.5
    h = -(x[:, None] ** 2 + x[None, :] ** 2) / (2 * R)
    fn = str(tmp_path / "sphere.txt")
    np.savetxt(fn, h)
    return fn, h


def test_pressure_programme_matches_direct_solves(gpu, tmp_path):
    fn, h = sphere_file(tmp_path)
    nc = str(tmp_path / "out.nc")
    pfn = str(tmp_path / "p.txt")
    rc = HardWall.main([fn, "--physical_sizes", "1,1", "--pressure", "0.001,0.003,3", "--pentol", "1e-10",
                        "--netcdf-fn", nc, "--pressure-fn", pfn, "--contact-fn", str(tmp_path / "c.txt"),
                        "--log-fn", str(tmp_path / "hw.out")], logger=quiet)
    assert rc == 0
    c = NetCDFContainer(nc)
    assert len(c) == 3
    system = make_system(PeriodicFFTElasticHalfSpace(h.shape, 1.0, (1.0, 1.0)), Topography(h, (1.0, 1.0)))
    for k, p in enumerate([0.001, 0.002, 0.003]):
        r = system.minimize_proxy(external_force=p, pentol=1e-10)
        np.testing.assert_allclose(c[k].forces, r.jac, rtol=0, atol=1e-12 * np.max(r.jac))
        assert c[k].load == pytest.approx(p)
        assert c[k].displacement == pytest.approx(r.offset)
        np.testing.assert_allclose(np.loadtxt(f"{pfn}.{k}"), r.jac / (1.0 / h.size), rtol=1e-9)
    assert (tmp_path / "c.txt.2").exists() and (tmp_path / "hw.out").exists()


def test_plastic_cone_like_reference(gpu, tmp_path):
    # test/commandline/test_hard_wall.py:18-56 (cone, E* 230000, H 6000)
    nx = ny = 128
    sx = sy = 0.005
    x = np.arange(nx).reshape(-1, 1) * sx / nx - sx / 2
    y = np.arange(ny).reshape(1, -1) * sy / ny - sy / 2
    fn = str(tmp_path / "cone.npy")
    np.save(fn, -np.sqrt(x ** 2 + y ** 2) * 0.05)
    nc = str(tmp_path / "cone.nc")
    rc = HardWall.main([fn, "--physical_sizes", f"{sx},{sy}", "--hardness", "6000", "--modulus", "230000",
                        "--pressure", f"{0.02 / (sx * sy)},{0.12 / (sx * sy)},3", "--boundary", "nonperiodic",
                        "--netcdf-fn", nc, "--log-fn", str(tmp_path / "hw.out")], logger=quiet)
    assert rc == 0
    c = NetCDFContainer(nc)
    assert len(c) == 3
    for k in range(3):
        assert c[k].forces.max() <= 6000 * (sx * sy / (nx * ny)) * (1 + 1e-9)


def test_displacement_and_automatic_programmes(gpu, tmp_path, monkeypatch):
    fn, h = sphere_file(tmp_path, n=32)
    rc = HardWall.main([fn, "--physical_sizes", "1,1", "--displacement=-0.02,0.0,3", "--log-fn",
                        str(tmp_path / "d.out"), "--netcdf-fn", str(tmp_path / "d.nc")], logger=quiet)
    assert rc == 0 and len(NetCDFContainer(str(tmp_path / "d.nc"))) == 3
    monkeypatch.setattr(HardWall, "NSTEPS", 6)
    rc = HardWall.main([fn, "--physical_sizes", "1,1", "--log-fn", str(tmp_path / "a.out"), "--netcdf-fn",
                        str(tmp_path / "a.nc"), "--gap-fn", str(tmp_path / "g.txt")], logger=quiet)
    assert rc == 0
    c = NetCDFContainer(str(tmp_path / "a.nc"))
    assert len(c) == 6
    gaps = [np.loadtxt(str(tmp_path / f"g.txt.{k}")) for k in range(6)]
    assert all(g.min() > -1e-6 for g in gaps)
