"""Grid, vertical metrics and namelist for the ARW dycore large step.

Mirrors the reference's ``DycoreMetrics`` (src/gpuwrf/contracts/grid.py:64-348,
flat constructor :166-245), the idealized pure-sigma override
(src/gpuwrf/ic_generators/idealized.py:379-451) and the hot-path namelist
fields of ``OperationalNamelist`` (src/gpuwrf/runtime/operational_mode.py:412-683).

All arrays are float64 numpy arrays in the reference's layout: 1-D vertical
metrics, mass-point map factors ``(ny, nx)``, u-point ``(ny, nx+1)`` and
v-point ``(ny+1, nx)``.
"""

from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class Metrics:
    """WRF-signed vertical metrics and map factors (contracts/grid.py:64-164)."""

    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    eta: np.ndarray  # (nz+1,) znw, 1 at the surface, 0 at the top
    p_top: float
    c1h: np.ndarray
    c2h: np.ndarray
    c1f: np.ndarray
    c2f: np.ndarray
    dnw: np.ndarray
    rdnw: np.ndarray
    dn: np.ndarray
    rdn: np.ndarray
    fnm: np.ndarray
    fnp: np.ndarray
    cf1: float
    cf2: float
    cf3: float
    msftx: np.ndarray
    msfty: np.ndarray
    msfux: np.ndarray
    msfuy: np.ndarray
    msfvx: np.ndarray
    msfvy: np.ndarray
    # dry-pressure coefficients p_d = c3 mu + c4 + p_top (read with
    # hypsometric_opt 2); None gives the idealized override c3 = 1, c4 = 0
    # (ic_generators/idealized.py:430-431)
    c3h: np.ndarray | None = None
    c4h: np.ndarray | None = None
    c3f: np.ndarray | None = None
    c4f: np.ndarray | None = None
    # Coriolis parameters on mass points (ny, nx) (contracts/grid.py:65-73):
    # f = 2 Omega sin(lat), e = 2 Omega cos(lat), and the projection rotation
    # sina/cosa; None gives no rotation (f = e = sina = 0, cosa = 1), the
    # idealized default under which every Coriolis term vanishes
    f: np.ndarray | None = None
    e: np.ndarray | None = None
    sina: np.ndarray | None = None
    cosa: np.ndarray | None = None

    def __post_init__(self) -> None:
        for name, n, v in (("c3h", self.nz, 1.0), ("c4h", self.nz, 0.0), ("c3f", self.nz + 1, 1.0),
                           ("c4f", self.nz + 1, 0.0)):
            if getattr(self, name) is None:
                object.__setattr__(self, name, np.full(n, v, dtype=np.float64))
        for name, v in (("f", 0.0), ("e", 0.0), ("sina", 0.0), ("cosa", 1.0)):
            a = getattr(self, name)
            if a is None:
                a = np.full((self.ny, self.nx), v, dtype=np.float64)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                if a.shape != (self.ny, self.nx):
                    # a no-rotation default carried over by dataclasses.replace to another grid
                    if np.all(a == v):
                        a = np.full((self.ny, self.nx), v, dtype=np.float64)
                    else:
                        raise ValueError(f"{name} must have the mass-point shape {(self.ny, self.nx)}, not {a.shape}")
            object.__setattr__(self, name, a)

    @property
    def rotating(self) -> bool:
        """True when any Coriolis term can be nonzero (f or e nonzero somewhere)."""

        return bool(np.any(self.f != 0.0) or np.any(self.e != 0.0))

    @classmethod
    def flat(cls, *, nx: int, ny: int, nz: int, dx: float, dy: float, eta: np.ndarray,
             p_top: float, pure_sigma: bool = True) -> "Metrics":
        """Flat unit-map-factor metrics, contracts/grid.py:166-245.

        ``pure_sigma`` applies the idealized override c1=1, c2=0
        (ic_generators/idealized.py:416-440).
        """

        if nz < 4:
            raise ValueError("the dycore needs nz >= 4 (flux3 stencils and cf coefficients)")
        eta = np.asarray(eta, dtype=np.float64)
        if eta.shape != (nz + 1,):
            raise ValueError("eta must have shape (nz+1,)")
        eta_mass = 0.5 * (eta[:-1] + eta[1:])
        dnw = eta[1:] - eta[:-1]
        rdnw = 1.0 / dnw
        dn = np.ones((nz,), dtype=np.float64)
        dn[1:] = 0.5 * (dnw[1:] + dnw[:-1])
        dn[0] = dnw[0]
        rdn = 1.0 / dn
        fnm = np.zeros((nz,), dtype=np.float64)
        fnp = np.zeros((nz,), dtype=np.float64)
        fnp[1:] = 0.5 * dnw[1:] / dn[1:]
        fnm[1:] = 0.5 * dnw[:-1] / dn[1:]
        cof1 = (2.0 * dn[1] + dn[2]) / (dn[1] + dn[2]) * dnw[0] / dn[1]
        cof2 = dn[1] / (dn[1] + dn[2]) * dnw[0] / dn[2]
        if pure_sigma:
            c1h = np.ones(nz)
            c2h = np.zeros(nz)
            c1f = np.ones(nz + 1)
            c2f = np.zeros(nz + 1)
            c3h, c3f = np.ones(nz), np.ones(nz + 1)
        else:
            c1h = eta_mass.copy()
            c2h = np.zeros(nz)
            c1f = eta.copy()
            c2f = np.zeros(nz + 1)
            c3h, c3f = eta_mass.copy(), eta.copy()
        one = lambda shape: np.ones(shape, dtype=np.float64)  # noqa: E731
        return cls(
            nx=int(nx), ny=int(ny), nz=int(nz), dx=float(dx), dy=float(dy), eta=eta.copy(),
            p_top=float(p_top), c1h=c1h, c2h=c2h, c1f=c1f, c2f=c2f, dnw=dnw, rdnw=rdnw, dn=dn,
            rdn=rdn, fnm=fnm, fnp=fnp, cf1=float(fnp[1] + cof1), cf2=float(fnm[1] - cof1 - cof2),
            cf3=float(cof2), msftx=one((ny, nx)), msfty=one((ny, nx)), msfux=one((ny, nx + 1)),
            msfuy=one((ny, nx + 1)), msfvx=one((ny + 1, nx)), msfvy=one((ny + 1, nx)),
            c3h=c3h, c4h=np.zeros(nz), c3f=c3f, c4f=np.zeros(nz + 1),
        )

    def shapes(self) -> dict[str, tuple[int, ...]]:
        """Shapes of the state fields, the moisture species included."""

        nx, ny, nz = self.nx, self.ny, self.nz
        return {
            **{q: (nz, ny, nx) for q in ("qv", "qc", "qr", "qi", "qs", "qg")},
            "u": (nz, ny, nx + 1),
            "v": (nz, ny + 1, nx),
            "w": (nz + 1, ny, nx),
            "theta": (nz, ny, nx),
            "p_total": (nz, ny, nx),
            "p_perturbation": (nz, ny, nx),
            "ph_total": (nz + 1, ny, nx),
            "ph_perturbation": (nz + 1, ny, nx),
            "mu_total": (ny, nx),
            "mu_perturbation": (ny, nx),
            "mudf": (ny, nx),
        }

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz


STATE_FIELDS = (
    "u", "v", "w", "theta", "p_total", "p_perturbation", "ph_total", "ph_perturbation",
    "mu_total", "mu_perturbation", "mudf",
)
# the moisture species a moist state adds, each (nz, ny, nx) (operational_mode.py:3100)
MOIST_FIELDS = ("qv", "qc", "qr", "qi", "qs", "qg")


def state_fields(namelist) -> tuple:
    """The state dict keys a step reads and a restart carries: the dry fields, and the six
    species when the namelist is moist."""

    return STATE_FIELDS + (MOIST_FIELDS if getattr(namelist, "moist", False) else ())


@dataclass(frozen=True)
class Namelist:
    """Hot-path namelist fields (runtime/operational_mode.py:412-683).

    Defaults follow the reference idealized gate
    (ic_generators/idealized.py:574-601): rigid lid, w_damping=1, damp_opt=3,
    dampcoef=0.2, flux-form advection, hypsometric_opt=1, force_fp64.
    """

    dt_s: float
    acoustic_substeps: int
    epssm: float = 0.1
    top_lid: bool = True
    w_damping: int = 1
    damp_opt: int = 3
    dampcoef: float = 0.2
    zdamp: float = 3000.0
    diff_6th_opt: int = 0
    diff_6th_factor: float = 0.12
    diff_opt: int = 0
    km_opt: int = 0
    c_s: float = 0.25
    scalar_adv_opt: int = 0
    const_nu_m2_s: float = 0.0
    # the state carries the six moisture species qv, qc, qr, qi, qs, qg
    # (the reference's State always does; a dry state leaves them out)
    moist: bool = False
    # moisture transport (operational_mode.py:495-522): 0 off, 1 positive-definite,
    # 2 monotonic limiter on the final RK stage, plain h5/v3 on the others
    moist_adv_opt: int = 0
    use_flux_advection: bool = True
    hypsometric_opt: int = 1
    h_sca_adv_order: int = 2
    force_fp64: bool = True
    disable_guards: bool = True
    run_physics: bool = False
    run_boundary: bool = False

    def validate(self) -> None:
        """Reject namelists outside the supported hot path (ValueError, as the reference)."""

        if self.acoustic_substeps < 1:
            raise ValueError("acoustic_substeps must be >= 1")
        if not self.top_lid:
            raise ValueError("only top_lid=True (rigid lid) is implemented")
        if not self.use_flux_advection:
            raise ValueError("only use_flux_advection=True is implemented")
        if self.hypsometric_opt not in (1, 2):
            raise ValueError("hypsometric_opt must be 1 (linear) or 2 (log-pressure)")
        if self.h_sca_adv_order > 2:
            raise ValueError("only the order-2 rhs_ph horizontal advection (h_sca_adv_order<=2) is implemented")
        if self.w_damping not in (0, 1):
            raise ValueError("w_damping must be 0 or 1")
        if self.damp_opt not in (0, 3):
            raise ValueError("damp_opt must be 0 or 3")
        if self.diff_6th_opt not in (0, 1, 2):
            raise ValueError("diff_6th_opt must be 0, 1 or 2")
        if self.diff_opt not in (0, 1) or self.km_opt not in (0, 4) or (self.diff_opt == 1 and self.km_opt != 4):
            raise ValueError("diffusion: only diff_opt=0, or diff_opt=1 with km_opt=4 (2-D Smagorinsky), is implemented")
        if not self.c_s >= 0.0:
            raise ValueError("c_s must be >= 0")
        if not (self.const_nu_m2_s >= 0.0 and math.isfinite(self.const_nu_m2_s)):
            raise ValueError("const_nu_m2_s must be finite and >= 0")
        if self.moist_adv_opt not in (0, 1, 2):
            raise ValueError("moist_adv_opt must be 0 (off), 1 (positive-definite) or 2 (monotonic)")
        if self.moist_adv_opt != 0 and not self.moist:
            raise ValueError("moist_adv_opt transports moisture: it needs moist=True (a state with qv..qg)")
        if self.const_nu_m2_s > 0.0 and self.diff_opt == 1:
            raise ValueError("const_nu_m2_s > 0 and the 2-D Smagorinsky diffusion do not run together")
        if self.scalar_adv_opt not in (0, 1, 2):
            raise ValueError("scalar_adv_opt must be 0, 1 (positive-definite) or 2 (monotonic)")
        if self.run_physics or self.run_boundary:
            raise ValueError("physics and lateral boundaries are outside the dycore hot path")

    def replace(self, **kw) -> "Namelist":
        return dataclasses.replace(self, **kw)
