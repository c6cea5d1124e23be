"""Classic netCDF files (openswpc_amd.cdf), checked against an independent
implementation of the format: scipy.io.netcdf_file reads what the writer
writes and writes what the reader reads."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from openswpc_amd import cdf


def _write_sample(path, version):
    w = cdf.Writer(path, version)
    w.dim("x", 5)
    w.dim("z", 3)
    w.dim("t", None)
    w.var("x", "f4", ["x"], {"long_name": "x", "units": "km"})
    w.var("z", "f4", ["z"])
    w.var("t", "f4", ["t"], {"units": "s"})
    w.var("rho", "f4", ["z", "x"], {"long_name": "rho"})
    w.var("Vx", "f4", ["t", "z", "x"], {"long_name": "Vx", "actual_range": np.float32([0, 0])})
    w.var("Vz", "f4", ["t", "z", "x"])
    w.var("n", "i4", ["x"])
    w.var("d", "f8", ["z"])
    w.gatt("title", "cdf test")
    w.gatt("ns1", 5)
    w.gatt("dt", np.float32(0.05))
    w.gatt("pair", np.float64([1.5, -2.0]))
    w.enddef()
    w.put("x", np.arange(5, dtype=np.float32) * 0.5)
    w.put("z", np.float32([1, 2, 3]))
    w.put("rho", np.arange(15, dtype=np.float32).reshape(3, 5))
    w.put("n", np.int32([1, -2, 3, -4, 5]))
    w.put("d", np.float64([0.1, 0.2, 0.3]))
    # records 0 and 2 written, record 1 left to the fill value
    for r in (0, 2):
        w.put_record("t", r, np.float32([r * 0.1]))
        w.put_record("Vx", r, np.full((3, 5), r + 1.0, np.float32))
        w.put_record("Vz", r, np.full((3, 5), -(r + 1.0), np.float32))
    w.att("Vx", "actual_range", np.float32([1, 3]))
    w.close()


@pytest.mark.parametrize("version", [1, 2])
def test_scipy_reads_writer_output(tmp_path, version):
    p = tmp_path / "a.nc"
    _write_sample(str(p), version)
    with netcdf_file(str(p), "r", mmap=False) as f:
        assert f.version_byte == version
        assert f.dimensions["x"] == 5 and f.dimensions["t"] is None
        assert f.title == b"cdf test"
        assert int(f.ns1) == 5
        assert np.float32(f.dt) == np.float32(0.05)
        np.testing.assert_array_equal(f.pair, [1.5, -2.0])
        np.testing.assert_array_equal(f.variables["x"][:], np.arange(5) * 0.5)
        np.testing.assert_array_equal(f.variables["rho"][:], np.arange(15).reshape(3, 5))
        np.testing.assert_array_equal(f.variables["n"][:], [1, -2, 3, -4, 5])
        np.testing.assert_array_equal(f.variables["d"][:], [0.1, 0.2, 0.3])
        vx = f.variables["Vx"][:]
        assert vx.shape == (3, 3, 5)
        assert np.all(vx[0] == 1) and np.all(vx[2] == 3)
        assert np.all(vx[1] == np.float32(9.9692099683868690e36))
        assert np.all(f.variables["Vz"][2] == -3)
        np.testing.assert_array_equal(f.variables["Vx"].actual_range, [1, 3])
        assert f.variables["x"].units == b"km"
        np.testing.assert_array_equal(f.variables["t"][:], np.float32([0, 9.9692099683868690e36, 0.2]))


def test_reader_reads_scipy_output(tmp_path):
    p = tmp_path / "b.nc"
    with netcdf_file(str(p), "w") as f:
        f.createDimension("x", 4)
        f.createDimension("y", 2)
        f.createDimension("z", 3)
        v = f.createVariable("x", "f4", ("x",))
        v[:] = [1, 2, 3, 4]
        f.createVariable("y", "f4", ("y",))[:] = [5, 6]
        f.createVariable("z", "f4", ("z",))[:] = [7, 8, 9]
        r = f.createVariable("rmed", "f4", ("z", "y", "x"))
        r[:] = np.arange(24, dtype=np.float32).reshape(3, 2, 4)
        r.long_name = "xi"
        f.title = "random media"
        f.epsilon = np.float32(0.05)
    ds = cdf.read(str(p))
    assert ds.dims == {"x": 4, "y": 2, "z": 3}
    assert ds.gatts["title"] == "random media"
    assert np.float32(ds.gatts["epsilon"][0]) == np.float32(0.05)
    np.testing.assert_array_equal(ds.variables["rmed"].data(), np.arange(24).reshape(3, 2, 4))
    assert ds.variables["rmed"].attrs["long_name"] == "xi"


def test_round_trip_records(tmp_path):
    p = tmp_path / "c.nc"
    _write_sample(str(p), 1)
    ds = cdf.read(str(p))
    assert ds.numrecs == 3 and ds.record_dim == "t"
    vx = ds.variables["Vx"].data()
    assert vx.shape == (3, 3, 5) and np.all(vx[2] == 3)
    assert cdf.file_size_ok(str(p))


def test_single_record_variable_unpadded(tmp_path):
    # a lone record variable of 2-byte values is not padded between records
    p = tmp_path / "d.nc"
    w = cdf.Writer(str(p))
    w.dim("t", None)
    w.dim("k", 3)
    w.var("s", "i2", ["t", "k"])
    w.enddef()
    for r in range(4):
        w.put_record("s", r, np.int16([r, r + 1, r + 2]))
    w.close()
    with netcdf_file(str(p), "r", mmap=False) as f:
        np.testing.assert_array_equal(f.variables["s"][:], [[r, r + 1, r + 2] for r in range(4)])
    assert cdf.read(str(p)).variables["s"].data()[3].tolist() == [3, 4, 5]


def test_attribute_change_after_enddef_must_keep_size(tmp_path):
    w = cdf.Writer(str(tmp_path / "e.nc"))
    w.dim("x", 2)
    w.var("x", "f4", ["x"], {"actual_range": np.float32([0, 0])})
    w.enddef()
    w.att("x", "actual_range", np.float32([-1, 1]))
    with pytest.raises(cdf.CDFError):
        w.att("x", "actual_range", np.float32([-1, 1, 2]))
    w.close()
