"""Host API of the MI355X dycore large step.

Mirrors the reference's step call: a state (u, v, w, theta, p, ph, mu and the
carried mudf) plus namelist parameters go in, the advanced state comes out --
``_physics_boundary_step(carry, namelist, step_index)`` with physics and
boundaries off (src/gpuwrf/runtime/operational_mode.py:4996-5013), whose core
is ``_rk_scan_step`` (:3319-3534).  The reference returns new arrays
(functional JAX, donating the state at its forecast entry,
operational_mode.py:5584); here the caller's device tensors are advanced in
place and returned, which is the donated-buffer behaviour.

Errors follow the reference: an unsupported namelist or mismatched state
raises ``ValueError`` before any device work; nothing is raised from inside
the step (NaNs propagate, as in the reference with guards off).

PyTorch only provides device memory and streams; all arithmetic runs in the
HIP kernels of libwrfdyn.so.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .grid import MOIST_FIELDS, STATE_FIELDS, Metrics, Namelist

# Reference State field names -> C-ABI wd_state member.
_ABI_FIELDS = (
    ("u", "u"), ("v", "v"), ("w", "w"), ("theta", "theta"), ("p_total", "p_total"),
    ("p_perturbation", "p_pert"), ("ph_total", "ph_total"), ("ph_perturbation", "ph_pert"),
    ("mu_total", "mu_total"), ("mu_perturbation", "mu_pert"), ("mudf", "mudf"),
)


def _c_array(a: np.ndarray, keep: list):
    a = np.ascontiguousarray(a, dtype=np.float64)
    keep.append(a)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _abi_structs(metrics: Metrics, namelist: Namelist, keep: list):
    """The C-ABI wd_grid and wd_namelist of a (grid, namelist); ``keep`` holds
    the host arrays the grid points to."""

    m = metrics
    grid = _lib.WdGrid(
        m.nx, m.ny, m.nz, m.dx, m.dy,
        _c_array(m.c1h, keep), _c_array(m.c2h, keep), _c_array(m.c1f, keep), _c_array(m.c2f, keep),
        _c_array(m.dnw, keep), _c_array(m.rdnw, keep), _c_array(m.dn, keep), _c_array(m.rdn, keep),
        _c_array(m.fnm, keep), _c_array(m.fnp, keep), m.cf1, m.cf2, m.cf3,
        _c_array(m.msftx, keep), _c_array(m.msfty, keep), _c_array(m.msfux, keep),
        _c_array(m.msfuy, keep), _c_array(m.msfvx, keep), _c_array(m.msfvy, keep),
        _c_array(m.c3h, keep), _c_array(m.c4h, keep), _c_array(m.c3f, keep), _c_array(m.c4f, keep),
        float(m.p_top),
        *((_c_array(m.f, keep), _c_array(m.e, keep), _c_array(m.sina, keep), _c_array(m.cosa, keep))
          if m.rotating else (None,) * 4),
    )
    nl = namelist
    nls = _lib.WdNamelist(float(nl.dt_s), int(nl.acoustic_substeps), float(nl.epssm),
                          int(bool(nl.top_lid)), int(nl.w_damping), int(nl.damp_opt),
                          float(nl.dampcoef), float(nl.zdamp), int(nl.diff_6th_opt),
                          float(nl.diff_6th_factor), int(nl.diff_opt), int(nl.km_opt),
                          float(nl.c_s), int(nl.scalar_adv_opt), int(nl.hypsometric_opt),
                          float(nl.const_nu_m2_s), int(bool(nl.moist)), int(nl.moist_adv_opt))
    return grid, nls


def _state_fields(namelist: Namelist) -> tuple:
    """The state dict keys a step reads: the dry fields, plus the six species when moist."""

    return STATE_FIELDS + (MOIST_FIELDS if namelist.moist else ())


def _check_tensors(state: dict, metrics: Metrics, namelist: Namelist) -> None:
    import torch

    shapes = metrics.shapes()
    fields = _state_fields(namelist)
    missing = [k for k in fields if k not in state]
    if missing:
        raise ValueError(f"state is missing fields {missing}")
    dev = None
    for name in fields:
        t = state[name]
        want = shapes.get(name, (metrics.nz, metrics.ny, metrics.nx))
        if not isinstance(t, torch.Tensor):
            raise ValueError(f"state[{name!r}] must be a torch tensor")
        if t.dtype != torch.float64:
            raise ValueError(f"state[{name!r}] must be float64 (force_fp64 path), got {t.dtype}")
        if tuple(t.shape) != want:
            raise ValueError(f"state[{name!r}] has shape {tuple(t.shape)}, expected {want}")
        if not t.is_cuda:
            raise ValueError(f"state[{name!r}] must live on the GPU")
        if not t.is_contiguous():
            raise ValueError(f"state[{name!r}] must be contiguous")
        dev = t.device if dev is None else dev
        if t.device != dev:
            raise ValueError("all state tensors must be on the same device")


class Dycore:
    """A compiled dycore for one (grid, namelist): the jit-cache entry analogue.

    ``step(state, n_steps)`` advances the state dict of float64 device
    tensors in place.  ``precision`` is ``"fp64"`` (the plain form,
    bit-identical to the CPU oracle) or ``"mixed"`` (FP32 advection
    interpolants and RK tendency storage, FP64 elsewhere; see
    ``wd_set_precision`` in include/wrfdyn.h).
    """

    PRECISIONS = {"fp64": _lib.WD_PREC_FP64, "mixed": _lib.WD_PREC_MIXED}

    def __init__(self, metrics: Metrics, namelist: Namelist, precision: str = "fp64"):
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(self.PRECISIONS)}, got {precision!r}")
        namelist.validate()
        self.metrics = metrics
        self.namelist = namelist
        lib = _lib.load()
        keep: list = []
        self._grid, self._nl = _abi_structs(metrics, namelist, keep)
        handle = ctypes.c_void_p()
        code = lib.wd_create(ctypes.byref(self._grid), ctypes.byref(self._nl), ctypes.byref(handle))
        if code != _lib.WD_OK:
            raise _lib.WrfdynError(code, _lib.last_error(None))
        self._handle = handle
        self._lib = lib
        if precision != "fp64":
            _lib.check(lib.wd_set_precision(handle, self.PRECISIONS[precision]), handle)
        self.precision = precision

    @property
    def workspace_bytes(self) -> int:
        return int(self._lib.wd_workspace_bytes(self._handle))

    def check_state(self, state: dict) -> None:
        """The fields a step reads (the six species too when the namelist is moist) are
        contiguous float64 device tensors of the grid's shapes, on one device."""

        _check_tensors(state, self.metrics, self.namelist)

    def step(self, state: dict, n_steps: int = 1, stream=None) -> dict:
        """Advance ``n_steps`` large steps in place on ``stream`` (default: torch's current)."""

        self.check_state(state)
        if n_steps < 0:
            raise ValueError("n_steps must be >= 0")
        stream = _stream_of(state, stream)
        ptrs = _state_ptrs(state, self.namelist.moist)
        code = self._lib.wd_step(self._handle, ctypes.byref(ptrs), int(n_steps),
                                 ctypes.c_void_p(stream.cuda_stream))
        if code != _lib.WD_OK:
            raise _lib.WrfdynError(code, _lib.last_error(self._handle))
        return state

    def step_adjoint(self, start: dict, bar: dict, stream=None) -> dict:
        """Adjoint of one large step from ``start`` (not modified): ``bar``
        holds the adjoint of the state after the step and is overwritten, in
        place, with the adjoint of ``start`` (every field an independent
        input).  FP64 form only.  The first call allocates the tape."""

        self.check_state(start)
        self.check_state(bar)
        stream = _stream_of(start, stream)
        a, b = _state_ptrs(start, self.namelist.moist), _state_ptrs(bar, self.namelist.moist)
        code = self._lib.wd_step_adjoint(self._handle, ctypes.byref(a), ctypes.byref(b),
                                         ctypes.c_void_p(stream.cuda_stream))
        if code != _lib.WD_OK:
            raise _lib.WrfdynError(code, _lib.last_error(self._handle))
        return bar

    @property
    def adjoint_tape_bytes(self) -> int:
        return int(self._lib.wd_adjoint_tape_bytes(self._handle))

    def set_precision(self, precision: str) -> None:
        """Switch the arithmetic form ("fp64" or "mixed") for later steps; drops a captured graph."""

        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(self.PRECISIONS)}, got {precision!r}")
        _lib.check(self._lib.wd_set_precision(self._handle, self.PRECISIONS[precision]), self._handle)
        self.precision = precision
        self._graph_state = None

    def capture_graph(self, state: dict) -> None:
        """Record one large step on ``state``'s buffers into a HIP graph (see graph_step).

        The tensors must stay alive, and keep their storage, while the graph is used.
        """

        self.check_state(state)
        ptrs = _state_ptrs(state, self.namelist.moist)
        _lib.check(self._lib.wd_capture_graph(self._handle, ctypes.byref(ptrs)), self._handle)
        self._graph_state = state

    def graph_step(self, n_steps: int = 1, stream=None) -> dict:
        """Replay the captured large step ``n_steps`` times; same results as ``step``."""

        state = getattr(self, "_graph_state", None)
        if state is None:
            raise ValueError("no captured graph: call capture_graph(state) first")
        if n_steps < 0:
            raise ValueError("n_steps must be >= 0")
        stream = _stream_of(state, stream)
        _lib.check(self._lib.wd_graph_step(self._handle, int(n_steps), ctypes.c_void_p(stream.cuda_stream)),
                   self._handle)
        return state

    def autotune(self, state: dict, trial_steps: int = 2, stream=None) -> dict:
        """Time the column-kernel forms and workgroup orders on a copy of ``state``; keep the fastest.

        Returns ``{"column_tile": t, "xcd_remap": r}`` (t = -1: column-march kernels).
        """

        self.check_state(state)
        stream = _stream_of(state, stream)
        ptrs = _state_ptrs(state, self.namelist.moist)
        _lib.check(self._lib.wd_autotune(self._handle, ctypes.byref(ptrs), int(trial_steps),
                                         ctypes.c_void_p(stream.cuda_stream)), self._handle)
        self._graph_state = None
        return self.tuned_config()

    def tuned_config(self) -> dict:
        t, r = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.wd_tuned_config(self._handle, ctypes.byref(t), ctypes.byref(r)), self._handle)
        return {"column_tile": t.value, "xcd_remap": r.value}

    def set_tuned_config(self, column_tile: int, xcd_remap: int) -> None:
        """Apply a tuning (as ``tuned_config`` returns it) without timing anything."""

        _lib.check(self._lib.wd_set_tuned_config(self._handle, int(column_tile), int(xcd_remap)), self._handle)
        self._graph_state = None

    def profile(self, kernels) -> None:
        """Time the named kernels (see _lib.kernel_names()) with HIP events in later steps."""

        names = _lib.kernel_names()
        mask = 0
        for k in kernels:
            mask |= 1 << names.index(k)
        _lib.check(self._lib.wd_profile_enable(self._handle, mask), self._handle)

    def profile_read(self) -> dict:
        """Summed milliseconds and launch counts per profiled kernel since the last read."""

        names = _lib.kernel_names()
        n = len(names)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        code = self._lib.wd_profile_read(self._handle, ms, cnt, n)
        if code < 0:
            raise _lib.WrfdynError(code, _lib.last_error(self._handle))
        return {names[i]: (ms[i], cnt[i]) for i in range(n) if cnt[i] > 0}

    def close(self) -> None:
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.wd_destroy(h)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Independent members with differing grids, advanced together.

    One large step of every member is captured into one HIP graph, the
    members spread over ``branches`` parallel branches; ``step`` launches it
    once per large step.  ``members`` is a list of (Dycore, state) pairs; the
    handles and tensors must outlive the ensemble.  Each member's result is
    identical to ``Dycore.step`` on that member alone.

    ``autotune=True`` first runs each member's autotuner (``Dycore.autotune``,
    on a copy of its state, ``trial_steps`` timed steps per candidate);
    ``tuned`` then lists the choices made alone.  Tuned alone, a member does
    not see the other members running beside it, so if some member picked an
    LDS column tile, the batched step is then timed on copies of the states
    with those picks and with every member on the column march (same bits),
    and the faster is kept (``batch_trial``).  Each member keeps its tuned
    workgroup order.

    ``persistent=True`` captures instead one launch of the persistent step
    per large step (``wd_batch_create_persistent``): every kernel body of
    every member runs from one launch of ``workgroups`` resident workgroups
    that meet at a grid barrier between phases.  With ``autotune=True`` a
    member tuned to the 16x16 LDS column tile runs that tile inside the
    launch (FP64, up to 63 levels; ``column_kernel`` tells), unless the
    persistent step with every member on the column march is faster: both
    are timed on copies of the states (``persistent_trial``) and the faster
    kept.  Members must share the
    precision form and substep count and use no explicit diffusion, limiter
    or moisture; nothing else may run on the GPU during a step.  ``status``
    synchronises and raises if a barrier wait ran out since the last check.
    """

    def __init__(self, members, branches: int = 4, autotune: bool = False, trial_steps: int = 1,
                 persistent: bool = False):
        if not members:
            raise ValueError("an ensemble needs at least one member")
        self.members = list(members)
        lib = _lib.load()
        dev = None
        for dyn, state in self.members:
            dyn.check_state(state)
            dev = state["u"].device if dev is None else dev
            if state["u"].device != dev:
                raise ValueError("all members must live on the same device")
        self.tuned = [dyn.autotune(state, trial_steps) for dyn, state in self.members] if autotune else None
        self.batch_trial = None
        if autotune and any(cfg["column_tile"] >= 0 for cfg in self.tuned):
            self.batch_trial = self._batch_trial(trial_steps, persistent, branches)
        n = len(self.members)
        handles = (ctypes.c_void_p * n)(*[d._handle for d, _ in self.members])
        states = (_lib.WdState * n)(*[_state_ptrs(s, d.namelist.moist) for d, s in self.members])
        out = ctypes.c_void_p()
        if persistent:
            code = lib.wd_batch_create_persistent(handles, states, n, ctypes.byref(out))
        else:
            code = lib.wd_batch_create(handles, states, n, int(branches), ctypes.byref(out))
        if code != _lib.WD_OK:
            raise _lib.WrfdynError(code, _lib.last_error(None))
        self._batch = out
        self._lib = lib
        self.device = dev
        self.persistent = bool(persistent)

    @property
    def persistent_trial(self):
        """The batch trial of a persistent ensemble (None for a graph ensemble or without a trial)."""

        return self.batch_trial if self.persistent else None

    def _batch_trial(self, trial_steps: int, persistent: bool, branches: int) -> dict:
        """Time the batched step (the persistent step, or the graph on ``branches`` branches) with the
        members' own tuned column tiles and with every member on the column march (same bits), on
        copies of the states, and keep the faster: each member was tuned alone, without the other
        members beside it or the persistent kernel's occupancy."""

        import torch

        clones = [(d, {k: v.clone() for k, v in s.items()}) for d, s in self.members]
        stream = torch.cuda.current_stream(self.members[0][1]["u"].device)

        def timed() -> float:
            ens = Ensemble(clones, branches=branches, persistent=persistent)
            try:
                ens.step(1, stream)
                ens.status(stream)
                t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0.record(stream)
                ens.step(max(int(trial_steps), 1), stream)
                t1.record(stream)
                ens.status(stream)
                return t0.elapsed_time(t1)
            finally:
                ens.close()

        tuned_ms = timed()
        for (d, _), cfg in zip(self.members, self.tuned):
            d.set_tuned_config(-1, cfg["xcd_remap"])
        march_ms = timed()
        kept = "tuned" if tuned_ms < march_ms else "march"
        if kept == "tuned":
            for (d, _), cfg in zip(self.members, self.tuned):
                d.set_tuned_config(cfg["column_tile"], cfg["xcd_remap"])
        del clones
        return {"tuned_ms": tuned_ms, "march_ms": march_ms, "kept": kept}

    @property
    def workgroups(self) -> int:
        """Workgroups of the persistent step's launch (0 for the graph of kernel launches)."""

        return int(self._lib.wd_batch_workgroups(self._batch))

    def column_kernel(self, member: int) -> int:
        """The acoustic column kernel ``member`` runs in the ensemble's step: -1 the column march, else a tile index."""

        out = ctypes.c_int(0)
        self._check(self._lib.wd_batch_column_kernel(self._batch, int(member), ctypes.byref(out)))
        return out.value

    def status(self, stream=None) -> None:
        """Synchronise ``stream``; raise if a persistent step's barrier wait ran out since the last call."""

        stream = self._stream(stream)
        self._check(self._lib.wd_batch_status(self._batch, ctypes.c_void_p(stream.cuda_stream)))

    def set_spin_limit(self, polls: int) -> None:
        """Bound one grid barrier wait of the persistent step to ``polls`` polls."""

        self._check(self._lib.wd_batch_set_spin_limit(self._batch, int(polls)))

    def branch(self, member: int) -> int:
        return int(self._lib.wd_batch_branch(self._batch, int(member)))

    @property
    def branches(self) -> int:
        return int(self._lib.wd_batch_branches(self._batch))

    def _check(self, code: int) -> None:
        if code != _lib.WD_OK:
            raise _lib.WrfdynError(code, self._lib.wd_batch_last_error(self._batch).decode())

    def step(self, n_steps: int = 1, stream=None) -> None:
        if n_steps < 0:
            raise ValueError("n_steps must be >= 0")
        stream = self._stream(stream)
        self._check(self._lib.wd_step_batched(self._batch, int(n_steps), ctypes.c_void_p(stream.cuda_stream)))

    def _stream(self, stream):
        import torch

        return torch.cuda.current_stream(self.device) if stream is None else stream

    @property
    def cells(self) -> int:
        return sum(d.metrics.nx * d.metrics.ny * d.metrics.nz for d, _ in self.members)

    def close(self) -> None:
        b = getattr(self, "_batch", None)
        if b is not None and b.value:
            self._lib.wd_batch_destroy(b)
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _state_ptrs(state: dict, moist: bool = False):
    """The wd_state of a state dict: the species' pointers only for a moist handle (NULL otherwise)."""

    ptrs = [ctypes.c_void_p(state[k].data_ptr()) for k, _ in _ABI_FIELDS]
    if moist:
        ptrs += [ctypes.c_void_p(state[k].data_ptr()) for k in MOIST_FIELDS]
    return _lib.WdState(*ptrs)


def _stream_of(state: dict, stream):
    import torch

    return torch.cuda.current_stream(state["u"].device) if stream is None else stream


def state_to_device(state: dict, device="cuda") -> dict:
    """Copy a dict of numpy arrays (the IC builders' output) to float64 device tensors."""

    import torch

    return {k: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)).to(device)
            for k, v in state.items()}


def state_to_host(state: dict) -> dict:
    return {k: v.detach().cpu().numpy() for k, v in state.items()}


def dycore_step(state: dict, namelist: Namelist, metrics: Metrics, n_steps: int = 1) -> dict:
    """One-shot drop-in for the reference step call: state dict in, advanced state out.

    For repeated calls build a :class:`Dycore` once (the reference's jit
    cache) and call ``step``.
    """

    import torch

    dyn = Dycore(metrics, namelist)
    out = dyn.step(state, n_steps)
    torch.cuda.synchronize(state["u"].device)
    dyn.close()
    return out


def step_noalloc(state: dict, namelist: Namelist, metrics: Metrics, n_steps: int = 1, stream=None) -> dict:
    """``n_steps`` large steps of ``state`` (float64 device tensors, advanced in
    place) through wd_step_noalloc: no handle and no device allocation, the
    library's fixed device scratch holds the metrics and the workspace.  Same
    results as :meth:`Dycore.step`; raises for a grid too large for the scratch."""

    namelist.validate()
    _check_tensors(state, metrics, namelist)
    lib = _lib.load()
    keep: list = []
    grid, nls = _abi_structs(metrics, namelist, keep)
    s = _state_ptrs(state, namelist.moist)
    stream = _stream_of(state, stream)
    code = lib.wd_step_noalloc(ctypes.byref(grid), ctypes.byref(nls), ctypes.byref(s), int(n_steps),
                               ctypes.c_void_p(stream.cuda_stream))
    if code != _lib.WD_OK:
        raise _lib.WrfdynError(code, _lib.last_error(None))
    return state
