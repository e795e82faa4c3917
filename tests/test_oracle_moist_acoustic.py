"""Pins of the oracle's two uses of the moisture factors inside the acoustic loop.

* advance_uv (dynamics/core/acoustic.py:720): u -= dts cqu dpx, and v the same
  with cqv: the factor scales the pressure gradient and nothing else.
* advance_w (core/advance_w.py:424-433): the explicit part termA of the
  implicit w equation carries cq1 = 1/(1 + cqw): the factor scales that term
  and nothing else.

These enter a whole step only at second order in dt (the work p' and phi'
they act on start each stage at zero), so they are pinned on the acoustic
pieces themselves (oracle.piece_moist), with inputs traced from a real step:

1. Where the term's inputs vanish (zero work p' and phi', uniform base
   pressure and php; for w also a zero phi-equation rhs), the piece's output
   is the same bits for any factors: the factor touches no other term (the
   large-step tendency, the external-mode damping, termB, rw_tend).
2. Elsewhere the output is affine in each face's own factor: with per-face
   factors c and a uniform c0, (x(1) - x(c)) / (x(1) - x(c0)) is
   (1 - c) / (1 - c0) at every face, so the factor multiplies one additive
   term at its own face (a factor read at a neighbouring face, or cqv on u
   faces, fails).  For w the tridiagonal solve is made the identity
   (ca = cgamma = 0, calpha = 1) so that faces do not mix.
"""

import numpy as np
import pytest

from oracle import oracle


def _before(piece_id, stage=3, substep=2):
    """A terrain case and the arrays before the given acoustic piece of its first large step."""

    from wrf_gpu_amd import ic

    case = ic.stratified_terrain_3d(nx=16, ny=12, nz=10, dx=3000.0, u_sigma=1.0, sigma=0.5)
    m, nl = case.metrics, case.namelist
    arrays = oracle.new_arrays(m, case.state)
    seen = {}
    for pc in oracle.step_pieces(nl):
        pid, rk, dts = pc
        if pid == piece_id and rk == stage:
            seen[pid] = seen.get(pid, 0) + 1
            if seen[pid] == substep:
                return case, arrays, rk, dts
        oracle.piece(m, nl, arrays, *pc)
    raise AssertionError("piece not reached")


def _factors(m, rng, raw_w=None):
    nx, ny, nz = m.nx, m.ny, m.nz
    cqu = np.ones((nz, ny, nx + 1))
    cqv = np.ones((nz, ny + 1, nx))
    cqwr = np.zeros((nz + 1, ny, nx)) if raw_w is None else raw_w
    return cqu, cqv, cqwr


def _run(case, arrays, piece_id, rk, dts, cqu, cqv, cqwr, nl=None):
    a = {k: v.copy() for k, v in arrays.items()}
    oracle.piece_moist(case.metrics, nl or case.namelist, a, piece_id, rk, dts, cqu, cqv, cqwr)
    return a


def _ratio_check(x1, xc, xc0, c, c0, what):
    d, d0 = x1 - xc, x1 - xc0
    mask = np.abs(d0) > 1e-9 * np.abs(d0).max()
    assert mask.sum() > 0.8 * mask.size, what  # the term is live almost everywhere
    got = d[mask] / d0[mask]
    want = ((1.0 - c) / (1.0 - c0))[mask]
    err = float(np.max(np.abs(got - want) / np.abs(want)))
    print(f"{what}: {mask.sum()} faces, max rel error of the per-face factor ratio {err:.1e}")
    assert err < 1e-8, (what, err)


def test_uv_moisture_factor_scales_only_the_pressure_gradient():
    case, arrays, rk, dts = _before(oracle.PIECE_UV)
    m = case.metrics
    assert np.abs(arrays["p"]).max() > 0 and np.abs(arrays["mudf"]).max() > 0 and np.abs(arrays["ut"]).max() > 0
    rng = np.random.default_rng(11)
    ones_u, ones_v, raw0 = _factors(m, rng)
    # 2. per-face affine
    cu = rng.uniform(0.80, 0.99, ones_u.shape)
    cv = rng.uniform(0.80, 0.99, ones_v.shape)
    c0 = 0.9
    x1 = _run(case, arrays, oracle.PIECE_UV, rk, dts, ones_u, ones_v, raw0)
    xc = _run(case, arrays, oracle.PIECE_UV, rk, dts, cu, cv, raw0)
    xc0 = _run(case, arrays, oracle.PIECE_UV, rk, dts, np.full_like(cu, c0), np.full_like(cv, c0), raw0)
    _ratio_check(x1["u_work"], xc["u_work"], xc0["u_work"], cu, c0, "u, cqu")
    _ratio_check(x1["v_work"], xc["v_work"], xc0["v_work"], cv, c0, "v, cqv")
    # the dry piece is the factor 1
    dry = {k: v.copy() for k, v in arrays.items()}
    oracle.piece(m, case.namelist, dry, oracle.PIECE_UV, rk, dts)
    assert np.array_equal(dry["u_work"], x1["u_work"]) and np.array_equal(dry["v_work"], x1["v_work"])
    # 1. no pressure gradient: no dependence on the factors, bit for bit
    flat = {k: v.copy() for k, v in arrays.items()}
    flat["p"][...] = 0.0
    flat["ph_work"][...] = 0.0
    flat["pb"][...] = flat["pb"].mean()
    flat["php"][...] = flat["php"].mean()
    y1 = _run(case, flat, oracle.PIECE_UV, rk, dts, ones_u, ones_v, raw0)
    yc = _run(case, flat, oracle.PIECE_UV, rk, dts, cu, cv, raw0)
    assert not np.array_equal(y1["u_work"], flat["u_work"])  # the tendency and the damping still act
    assert np.array_equal(y1["u_work"], yc["u_work"]) and np.array_equal(y1["v_work"], yc["v_work"])


@pytest.mark.parametrize("plant", [None, "neighbour"])
def test_w_moisture_factor_scales_only_term_a(plant):
    case, arrays, rk, dts = _before(oracle.PIECE_W)
    m, nx, ny, nz = case.metrics, case.metrics.nx, case.metrics.ny, case.metrics.nz
    nl = case.namelist.replace(damp_opt=0, w_damping=0)
    ident = {k: v.copy() for k, v in arrays.items()}
    ident["ca"][...] = 0.0
    ident["cgamma"][...] = 0.0
    ident["calpha"][...] = 1.0
    assert np.abs(ident["ph_work"]).max() > 0
    rng = np.random.default_rng(12)
    ones_u, ones_v, raw0 = _factors(m, rng)
    # cq1 = 1 / (1 + raw) on the interior faces
    c = rng.uniform(0.80, 0.99, (nz + 1, ny, nx))
    c0 = 0.9
    raw = 1.0 / c - 1.0
    raw0_c = np.full_like(raw, 1.0 / c0 - 1.0)
    x1 = _run(case, ident, oracle.PIECE_W, rk, dts, ones_u, ones_v, raw0, nl)
    xc = _run(case, ident, oracle.PIECE_W, rk, dts, ones_u, ones_v, raw, nl)
    xc0 = _run(case, ident, oracle.PIECE_W, rk, dts, ones_u, ones_v, raw0_c, nl)
    w1, wc, wc0 = (x["w_work"][1:nz] for x in (x1, xc, xc0))
    want_c = c[1:nz] if plant is None else np.roll(c, 1, axis=0)[1:nz]  # a factor one face down must fail
    if plant is None:
        _ratio_check(w1, wc, wc0, want_c, c0, "w, cq1")
    else:
        with pytest.raises(AssertionError):
            _ratio_check(w1, wc, wc0, want_c, c0, "w, cq1 of the face below")
    # 1. termA's inputs zero (work phi', the phi-equation rhs): no dependence on cq1, bit for bit
    flat = {k: v.copy() for k, v in ident.items()}
    for k in ("ph_work", "ph_tend", "w_work", "ww"):
        flat[k][...] = 0.0
    y1 = _run(case, flat, oracle.PIECE_W, rk, dts, ones_u, ones_v, raw0, nl)
    yc = _run(case, flat, oracle.PIECE_W, rk, dts, ones_u, ones_v, raw, nl)
    assert np.abs(y1["w_work"][1:nz]).max() > 0  # termB and rw_tend still act
    assert np.array_equal(y1["w_work"], yc["w_work"])
