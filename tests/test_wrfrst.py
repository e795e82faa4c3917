"""WRF-compatible netCDF restart of the dycore carry (wrf_gpu_amd.wrfrst,
mirroring src/gpuwrf/io/wrfrst_netcdf.py:574-728): the WRF layer carries the
standard names, dimensions and attributes with the values WRF defines
(T = theta - 300, PB = p_total - P, ...) to float32; the exact layer
round-trips bit for bit, so an oracle run resumed from it equals the
straight run; schema and layer mismatches are refused."""

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic, wrfrst
from wrf_gpu_amd.grid import STATE_FIELDS


@pytest.fixture
def case():
    c = ic.stratified_terrain_3d(nx=14, ny=11, nz=6, dx=3000.0, u_sigma=1.0, sigma=0.5)
    return c.__class__(**{**c.__dict__, "namelist": c.namelist.replace(diff_opt=1, km_opt=4)})


def test_wrf_layer_has_wrf_names_dims_and_values(case, tmp_path):
    st = oracle.run(case.metrics, case.namelist, case.state, 1)
    p = wrfrst.write_wrfrst(st, case.metrics, case.namelist, 1, tmp_path / "wrfrst_d01")
    sch = wrfrst.inspect_wrfrst_schema(p)
    v = sch["variables"]
    assert v["U"]["dimensions"] == ["Time", "bottom_top", "south_north", "west_east_stag"]
    assert v["V"]["dimensions"] == ["Time", "bottom_top", "south_north_stag", "west_east"]
    assert v["PH"]["dimensions"] == ["Time", "bottom_top_stag", "south_north", "west_east"]
    assert v["MU"]["dimensions"] == ["Time", "south_north", "west_east"]
    assert v["T"]["units"] == "K" and v["P"]["units"] == "Pa"
    g = sch["global_attrs"]
    assert g["WEST-EAST_GRID_DIMENSION"] == 15 and g["BOTTOM-TOP_GRID_DIMENSION"] == 7
    from scipy.io import netcdf_file

    with netcdf_file(str(p), "r", mmap=False) as f:
        assert f.variables["T"][0].dtype.kind == "f" and f.variables["T"][0].dtype.itemsize == 4
        assert np.array_equal(f.variables["T"][0], (st["theta"] - 300.0).astype(np.float32))
        assert np.array_equal(f.variables["PB"][0], (st["p_total"] - st["p_perturbation"]).astype(np.float32))
        assert np.array_equal(f.variables["MUB"][0], (st["mu_total"] - st["mu_perturbation"]).astype(np.float32))
        assert bytes(f.variables["Times"][0]).decode() == "0001-01-01_00:00:00"


def test_exact_layer_round_trips_and_resumes_bit_for_bit(case, tmp_path):
    straight = oracle.run(case.metrics, case.namelist, case.state, 3)
    one = oracle.run(case.metrics, case.namelist, case.state, 1)
    p = wrfrst.write_wrfrst(one, case.metrics, case.namelist, 1, tmp_path / "wrfrst_d01")
    st, m, nl, step = wrfrst.read_wrfrst(p)
    assert step == 1 and nl == case.namelist
    for k in STATE_FIELDS:
        assert np.array_equal(st[k], one[k]), k
    for f in ("c1h", "rdnw", "fnm", "msfux", "msfvy", "eta"):
        assert np.array_equal(getattr(m, f), getattr(case.metrics, f)), f
    resumed = oracle.run(m, nl, st, 3 - step)
    for k in STATE_FIELDS:
        assert np.array_equal(resumed[k], straight[k]), k


def test_disagreeing_layers_and_schema_are_refused(case, tmp_path):
    from scipy.io import netcdf_file

    p = wrfrst.write_wrfrst(case.state, case.metrics, case.namelist, 0, tmp_path / "wrfrst_d01")
    with netcdf_file(str(p), "a", mmap=False) as f:
        t = np.array(f.variables["T"][0])
        t[0, 0, 0] += 1.0
        f.variables["T"][0] = t
    with pytest.raises(ValueError, match="disagrees"):
        wrfrst.read_wrfrst(p)
    p2 = wrfrst.write_wrfrst(case.state, case.metrics, case.namelist, 0, tmp_path / "wrfrst_d02")
    with netcdf_file(str(p2), "a", mmap=False) as f:
        f.GPUWRF_STATE_FIELD_ORDER = "u,v"
    with pytest.raises(ValueError, match="field order"):
        wrfrst.read_wrfrst(p2)


def test_moist_restart_has_wrf_q_variables_and_resumes_bit_for_bit(case, tmp_path):
    from scipy.io import netcdf_file

    mcase = ic.with_moisture(case, moist_adv_opt=1)
    straight = oracle.run(mcase.metrics, mcase.namelist, mcase.state, 3)
    one = oracle.run(mcase.metrics, mcase.namelist, mcase.state, 1)
    p = wrfrst.write_wrfrst(one, mcase.metrics, mcase.namelist, 1, tmp_path / "wrfrst_d01")
    with netcdf_file(str(p), "r", mmap=False) as f:
        for name, key in (("QVAPOR", "qv"), ("QCLOUD", "qc"), ("QRAIN", "qr"), ("QICE", "qi"), ("QSNOW", "qs"),
                          ("QGRAUP", "qg")):
            v = f.variables[name]
            assert v.dimensions == ("Time", "bottom_top", "south_north", "west_east") and v.units == b"kg kg-1"
            assert np.array_equal(v[0], one[key].astype(np.float32)), name
    st, m, nl, step = wrfrst.read_wrfrst(p)
    assert nl == mcase.namelist
    resumed = oracle.run(m, nl, st, 3 - step)
    for k in STATE_FIELDS + ("qv", "qc", "qr", "qi", "qs", "qg"):
        assert np.array_equal(resumed[k], straight[k]), k
