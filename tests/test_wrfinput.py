"""Ingest of a WRF netCDF file (wrf_gpu_amd.wrfinput): a file written with
WRF's variable names, dimensions and perturbation/base split from a known
case gives back its metrics bit for bit and its state to one rounding of the
recombined totals when the file holds doubles (to float32 rounding when it
holds floats, as WRF writes them), so the oracle steps the ingested case as
the original to 1e-11; files that lack a variable the step needs are
refused."""

import numpy as np
import pytest

from oracle import oracle
from wrf_gpu_amd import ic, wrfinput, wrfrst
from wrf_gpu_amd.grid import STATE_FIELDS

_VERT = {"ZNW": "eta", "C1H": "c1h", "C2H": "c2h", "C1F": "c1f", "C2F": "c2f", "DNW": "dnw", "RDNW": "rdnw",
         "DN": "dn", "RDN": "rdn", "FNM": "fnm", "FNP": "fnp"}
_MAPS = {"MAPFAC_MX": "msftx", "MAPFAC_MY": "msfty", "MAPFAC_UX": "msfux", "MAPFAC_UY": "msfuy",
         "MAPFAC_VX": "msfvx", "MAPFAC_VY": "msfvy"}


def _write(path, case, dtype="f8", drop=None, extra=None, attrs=None):
    from scipy.io import netcdf_file

    m, s = case.metrics, case.state
    f = netcdf_file(str(path), "w", version=2)
    f.createDimension("Time", 1)
    for name, n in (("west_east", m.nx), ("south_north", m.ny), ("bottom_top", m.nz), ("west_east_stag", m.nx + 1),
                    ("south_north_stag", m.ny + 1), ("bottom_top_stag", m.nz + 1)):
        f.createDimension(name, n)

    def put(name, dims, a, t=dtype):
        if name == drop:
            return
        v = f.createVariable(name, t, dims)
        v[:] = np.asarray(a)[None] if dims and dims[0] == "Time" else a

    for name, k in _VERT.items():
        a = getattr(m, k)
        put(name, ("Time", "bottom_top_stag" if a.shape[0] == m.nz + 1 else "bottom_top"), a, "f8")
    for name, k in _MAPS.items():
        a = getattr(m, k)
        put(name, ("Time", "south_north_stag" if a.shape[0] == m.ny + 1 else "south_north",
                   "west_east_stag" if a.shape[1] == m.nx + 1 else "west_east"), a, "f8")
    for name, k in (("CF1", "cf1"), ("CF2", "cf2"), ("CF3", "cf3"), ("P_TOP", "p_top")):
        put(name, ("Time",), np.array([getattr(m, k)]).reshape(()), "f8")
    put("U", ("Time", "bottom_top", "south_north", "west_east_stag"), s["u"])
    put("V", ("Time", "bottom_top", "south_north_stag", "west_east"), s["v"])
    put("W", ("Time", "bottom_top_stag", "south_north", "west_east"), s["w"])
    put("T", ("Time", "bottom_top", "south_north", "west_east"), s["theta"] - 300.0)
    put("P", ("Time", "bottom_top", "south_north", "west_east"), s["p_perturbation"])
    put("PB", ("Time", "bottom_top", "south_north", "west_east"), s["p_total"] - s["p_perturbation"])
    put("PH", ("Time", "bottom_top_stag", "south_north", "west_east"), s["ph_perturbation"])
    put("PHB", ("Time", "bottom_top_stag", "south_north", "west_east"), s["ph_total"] - s["ph_perturbation"])
    put("MU", ("Time", "south_north", "west_east"), s["mu_perturbation"])
    put("MUB", ("Time", "south_north", "west_east"), s["mu_total"] - s["mu_perturbation"])
    for name, a in (extra or {}).items():
        put(name, ("Time", "south_north", "west_east"), a, "f8")
    f.DX = np.float32(m.dx)
    f.DY = np.float32(m.dy)
    for k, v in (attrs or {}).items():
        setattr(f, k, v)
    f.close()
    return path


@pytest.fixture
def case():
    return ic.stratified_terrain_3d(nx=14, ny=11, nz=6, dx=3000.0, u_sigma=1.0, sigma=0.5)


def test_double_file_gives_the_case_back(case, tmp_path):
    st, m = wrfinput.read_wrfinput(_write(tmp_path / "wrfinput_d01", case))
    for f in ("eta", "c1h", "c2f", "rdnw", "fnm", "msfuy", "msfvx"):
        assert np.array_equal(getattr(m, f), getattr(case.metrics, f)), f
    assert (m.nx, m.ny, m.nz, m.dx, m.dy, m.cf1, m.p_top) == (case.metrics.nx, case.metrics.ny, case.metrics.nz,
                                                                case.metrics.dx, case.metrics.dy, case.metrics.cf1,
                                                                case.metrics.p_top)
    # the perturbation/base split round-trips to within one rounding of the totals
    for k in STATE_FIELDS:
        scale = max(np.abs(case.state[k]).max(), 1.0)
        assert np.allclose(st[k], case.state[k], rtol=0, atol=2e-16 * scale * 4), k
    a = oracle.run(case.metrics, case.namelist, case.state, 2)
    b = oracle.run(m, case.namelist, st, 2)
    for k in STATE_FIELDS:
        scale = max(np.abs(a[k]).max(), 1.0)
        assert np.allclose(a[k], b[k], rtol=0, atol=1e-11 * scale), k


def test_float_file_as_wrf_writes_it(case, tmp_path):
    st, _ = wrfinput.read_wrfinput(_write(tmp_path / "wrfinput_d01", case, dtype="f4"))
    assert np.array_equal(st["u"], case.state["u"].astype(np.float32).astype(np.float64))
    assert np.array_equal(st["mu_perturbation"], case.state["mu_perturbation"].astype(np.float32).astype(np.float64))
    assert np.allclose(st["theta"], case.state["theta"], rtol=1e-6)


def test_a_file_without_the_vertical_coordinate_is_refused(case, tmp_path):
    # this build's wrfrst keeps its metrics in the exact layer only
    p = wrfrst.write_wrfrst(case.state, case.metrics, case.namelist, 0, tmp_path / "wrfrst_d01")
    with pytest.raises(ValueError, match="ZNW"):
        wrfinput.read_wrfinput(p)


@pytest.mark.parametrize("drop", ["T", "MAPFAC_UX", "C1H"])
def test_missing_variables_are_refused(case, tmp_path, drop):
    with pytest.raises(ValueError, match=drop):
        wrfinput.read_wrfinput(_write(tmp_path / "wrfinput_d01", case, drop=drop))


def test_coriolis_parameters_are_read(case, tmp_path):
    ny, nx = case.metrics.ny, case.metrics.nx
    rng = np.random.default_rng(3)
    F, E = 1e-4 + 1e-6 * rng.random((ny, nx)), 5e-5 + 1e-6 * rng.random((ny, nx))
    a = 0.1 * rng.random((ny, nx))
    p = _write(tmp_path / "wrfinput_d01", case, extra={"F": F, "E": E, "SINALPHA": np.sin(a), "COSALPHA": np.cos(a)},
               attrs={"MAP_PROJ": np.int32(0)})
    _, m = wrfinput.read_wrfinput(p)
    assert m.rotating
    for k, want in (("f", F), ("e", E), ("sina", np.sin(a)), ("cosa", np.cos(a))):
        assert np.array_equal(getattr(m, k), want), k
    # without F and E the reference's fallback: 2 Omega sin / cos of XLAT
    lat = 30.0 + rng.random((ny, nx))
    _, m = wrfinput.read_wrfinput(_write(tmp_path / "wrfinput_d02", case, extra={"XLAT": lat}))
    assert np.allclose(m.f, 2 * 7.2921e-5 * np.sin(np.deg2rad(lat)), rtol=1e-15)
    assert np.allclose(m.e, 2 * 7.2921e-5 * np.cos(np.deg2rad(lat)), rtol=1e-15)
    assert not np.any(m.sina) and np.all(m.cosa == 1.0)
    # no rotation data at all: no rotation
    _, m = wrfinput.read_wrfinput(_write(tmp_path / "wrfinput_d03", case))
    assert not m.rotating


def test_a_real_data_file_is_refused_unless_stepped_as_periodic(case, tmp_path):
    p = _write(tmp_path / "wrfinput_d01", case, attrs={"MAP_PROJ": np.int32(1)})
    with pytest.raises(ValueError, match="MAP_PROJ"):
        wrfinput.read_wrfinput(p)
    st, m = wrfinput.read_wrfinput(p, assume_periodic=True)
    assert st["u"].shape == case.state["u"].shape


def test_moisture_species_are_read(case, tmp_path):
    """QVAPOR, QCLOUD and QRAIN present (a warm-rain microphysics): the state carries all
    six species, the absent ice species as zeros; without QVAPOR it carries none."""

    from scipy.io import netcdf_file

    mcase = ic.with_moisture(case)
    p = _write(tmp_path / "wrfinput_d01", case)
    with netcdf_file(str(p), "a") as f:
        for name, key in (("QVAPOR", "qv"), ("QCLOUD", "qc"), ("QRAIN", "qr")):
            v = f.createVariable(name, "f8", ("Time", "bottom_top", "south_north", "west_east"))
            v[:] = mcase.state[key][None]
    st, _ = wrfinput.read_wrfinput(p)
    for k in ("qv", "qc", "qr"):
        assert np.array_equal(st[k], mcase.state[k]), k
    for k in ("qi", "qs", "qg"):
        assert st[k].shape == mcase.state[k].shape and not np.any(st[k]), k
    st, _ = wrfinput.read_wrfinput(_write(tmp_path / "wrfinput_d02", case))
    assert "qv" not in st
