"""Differentiable smoke rollout: the hand-written adjoint of Part 5 of the C ABI behind a
torch.autograd.Function.

A rollout runs `nsteps` device-resident steps (force + attenuation with a per-step cell-centred
control force, velocity advection, mixed-precision projection, scalar advection, reinjection) and
returns the final density.  Its backward recomputes each checkpoint segment and runs the adjoint
kernels; the gradient flows to the control forces (nsteps, D, H, W, 3).

torch provides device memory and the stream only; every operation runs in libzq_smoke3d.so.
"""
import ctypes as C

import torch

from . import _lib
from ._lib import check
from .sim import default_config

_FIELDS = {"u": _lib.FIELD_U, "v": _lib.FIELD_V, "w": _lib.FIELD_W, "T": _lib.FIELD_T, "rho": _lib.FIELD_RHO,
           "occ": _lib.FIELD_OCC, "p64": _lib.FIELD_P64}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class SmokeRollout:
    """Checkpointed rollout of `nsteps` steps on a W x H x D grid.

    dtype torch.float32 runs the product kernels; torch.float64 runs every field in FP64 (the
    finite-difference check mode).  Keyword overrides apply to the device config as in
    SmokeSim3D ("sor_" prefix for the projection options); mixed precision is switched on.
    """

    def __init__(self, W, H, D, nsteps, checkpoint_every, dtype=torch.float32, device="cuda", **overrides):
        if dtype not in (torch.float32, torch.float64):
            raise TypeError("dtype must be torch.float32 or torch.float64")
        cfg = default_config(W, H, D)
        cfg.sor.mixed_precision = 1
        for k, val in overrides.items():
            if k.startswith("sor_"):
                setattr(cfg.sor, k[4:], val)
            else:
                setattr(cfg, k, val)
        self.cfg = cfg
        self.W, self.H, self.D = W, H, D
        self.nsteps = nsteps
        self.dtype = dtype
        self.device = torch.device(device)
        h = C.c_void_p()
        prec = 4 if dtype == torch.float32 else 8
        check(_lib.lib().zq3d_adj_create(C.byref(cfg), prec, nsteps, checkpoint_every, C.byref(h)),
              "zq3d_adj_create")
        self._h = h
        self.last_forward = None
        self.last_backward = None

    def close(self):
        if self._h:
            torch.cuda.synchronize(self.device)
            _lib.lib().zq3d_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up ------------------------------------------------------------------
    def shape(self, name):
        W, H, D = self.W, self.H, self.D
        return {"u": (D, H, W + 1), "v": (D, H + 1, W), "w": (D + 1, H, W), "T": (D, H, W), "rho": (D, H, W),
                "occ": (D, H, W), "p64": (D + 2, H + 2, W + 2)}[name]

    def _stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(_lib.lib().zq3d_adj_set_stream(self._h, C.c_void_p(s)), "zq3d_adj_set_stream")

    def set_sources(self, boxes=(), spheres=()):
        b = torch.tensor(list(boxes), dtype=torch.float32).reshape(-1).contiguous()
        s = torch.tensor(list(spheres), dtype=torch.float32).reshape(-1).contiguous()
        check(_lib.lib().zq3d_adj_set_sources(self._h, b.numel() // 8, C.c_void_p(b.data_ptr()) if b.numel() else None,
                                              s.numel() // 6, C.c_void_p(s.data_ptr()) if s.numel() else None),
              "zq3d_adj_set_sources")

    def set_initial(self, **fields):
        """u, v, w, T, rho (value dtype), occ (uint8) and p64 (float64, the padded warm start of the
        forward projection), any subset; unset fields stay zero."""
        self._stream()
        for name, t in fields.items():
            want = torch.uint8 if name == "occ" else (torch.float64 if name == "p64" else self.dtype)
            t = torch.as_tensor(t).to(device=self.device, dtype=want).contiguous()
            if tuple(t.shape) != self.shape(name):
                raise ValueError("%s: shape %s, expected %s" % (name, tuple(t.shape), self.shape(name)))
            check(_lib.lib().zq3d_adj_set_initial(self._h, _FIELDS[name], _ptr(t)), "zq3d_adj_set_initial")
        torch.cuda.current_stream(self.device).synchronize()

    def set_adjoint_tol(self, rtol):
        """Adjoint projection tolerance relative to max|rhs|; 0 matches each step's forward projection."""
        check(_lib.lib().zq3d_adj_set_adjoint_tol(self._h, float(rtol)), "zq3d_adj_set_adjoint_tol")

    def set_pressure_store(self, on=True):
        """Keep each forward step's final FP64 pressure (nsteps padded grids) so that the backward's
        recomputed steps apply it instead of solving again; same states and gradient."""
        check(_lib.lib().zq3d_adj_set_pressure_store(self._h, int(bool(on))), "zq3d_adj_set_pressure_store")

    def set_sor_shape(self, ty, zc):
        """Sweep algorithm of every projection (a SmokeSim3D.autotune_sor candidate); same bits."""
        check(_lib.lib().zq3d_adj_set_sor_shape(self._h, int(ty), int(zc)), "zq3d_adj_set_sor_shape")

    # ---- rollout -------------------------------------------------------------------
    def controls_shape(self):
        return (self.nsteps, self.D, self.H, self.W, 3)

    def _check_controls(self, controls):
        if controls is None:
            return None
        if tuple(controls.shape) != self.controls_shape() or controls.dtype != self.dtype or not controls.is_cuda:
            raise ValueError("controls must be a %s cuda tensor of shape %s" % (self.dtype, self.controls_shape()))
        return controls.contiguous()

    def forward(self, controls=None):
        """Final density (D, H, W)."""
        controls = self._check_controls(controls)
        self._stream()
        rho = torch.empty((self.D, self.H, self.W), dtype=self.dtype, device=self.device)
        st = _lib.AdjStats()
        check(_lib.lib().zq3d_adj_forward(self._h, _ptr(controls), C.byref(st), _ptr(rho)), "zq3d_adj_forward")
        self.last_forward = st.as_dict()
        return rho

    def backward(self, controls, rho_bar):
        """dL/d(controls) from dL/d(final density)."""
        controls = self._check_controls(controls)
        rho_bar = rho_bar.to(dtype=self.dtype).contiguous()
        self._stream()
        g = torch.empty(self.controls_shape(), dtype=self.dtype, device=self.device)
        st = _lib.AdjStats()
        check(_lib.lib().zq3d_adj_backward(self._h, _ptr(controls), _ptr(rho_bar), _ptr(g), C.byref(st)),
              "zq3d_adj_backward")
        self.last_backward = st.as_dict()
        return g

    def evaluate(self, controls, from_step=0):
        """Final density for `controls`, rerun from the latest forward's checkpoint at or before
        `from_step` (the controls before that checkpoint must be the forward's); the forward's
        checkpoints, tape and pressure store are kept, so its backward is unaffected."""
        controls = self._check_controls(controls)
        self._stream()
        rho = torch.empty((self.D, self.H, self.W), dtype=self.dtype, device=self.device)
        st = _lib.AdjStats()
        check(_lib.lib().zq3d_adj_evaluate(self._h, int(from_step), _ptr(controls), C.byref(st), _ptr(rho)),
              "zq3d_adj_evaluate")
        return rho

    def generation(self):
        """Number of forwards run so far; a backward differentiates the latest one."""
        return int(_lib.lib().zq3d_adj_generation(self._h))

    def tape(self, step, name):
        """Recorded state of `step` in the segment on tape: u/v/w the projected faces of that step,
        T/rho the fields after its force stage."""
        out = torch.empty(self.shape(name), dtype=self.dtype, device=self.device)
        self._stream()
        check(_lib.lib().zq3d_adj_get_tape(self._h, int(step), _FIELDS[name], _ptr(out)), "zq3d_adj_get_tape")
        return out

    def get(self, name, grad=False):
        """Final state field after a forward, or dL/d(initial field) after a backward (grad=True)."""
        want = torch.uint8 if name == "occ" else self.dtype
        out = torch.empty(self.shape(name), dtype=want, device=self.device)
        self._stream()
        check(_lib.lib().zq3d_adj_get(self._h, _FIELDS[name], 1 if grad else 0, _ptr(out)), "zq3d_adj_get")
        return out


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, controls, engine):
        rho = engine.forward(controls)
        ctx.engine = engine
        ctx.generation = engine.generation()
        ctx.save_for_backward(controls)
        return rho

    @staticmethod
    def backward(ctx, grad_rho):
        (controls,) = ctx.saved_tensors
        if ctx.engine.generation() != ctx.generation:
            # the engine's checkpoints, tape and pressure store now belong to a later forward
            raise RuntimeError("rollout: the engine ran another forward after this one; its gradient can no "
                               "longer be computed (use one engine per rollout in a graph, or backward first)")
        return ctx.engine.backward(controls, grad_rho), None


def rollout(engine, controls):
    """Final density of `engine` driven by `controls` (nsteps, D, H, W, 3), differentiable in them."""
    return _Rollout.apply(controls, engine)
