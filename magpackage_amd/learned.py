"""Learned reluctivity predictor: warm start for the transient Newton solve.

No reference counterpart (MagPackage has no learned component).  For every
element of a nonlinear (B-H) region, a small MLP (include/mpk_mlp.h: 16 ->
64 -> 64 -> 64 -> 1, tanh, BF16 weights on the matrix cores, FP32
accumulation) maps the element's recent flux history

    B1, B2, B3    converged B of the last three steps (B_{n-1}, B_{n-2}, B_{n-3})
    Be  = 2 B1 - B2            (linear extrapolation to step n)
    dB  = B1 - (2 B2 - B3)     (how far the last step moved off its own
                                extrapolation)
    Bq  = 3 B1 - 3 B2 + B3     (quadratic extrapolation)

as the 16 inputs (B1, Be, 10 dB, |B1|^2, |Be|^2, drive(t_n), 100 |dB|^2,
|Bq|^2, region code; B in units of 2.2 T, |B|^2 of 4.84 T^2) to the
log-reluctivity the converged solution of step n will have, learnt as a
correction to log nu(|Be|^2) (the B-H curve at the extrapolated field);
``drive`` is an optional pair of global features of the prescribed source at
the new time (for a sinusoidal winding: sin and cos of its phase).
The warm start assembles K(nu_hat) + M_sigma/dt + reg M_reg (the Newton
system's secant form, whose solution is the converged field when nu_hat is
exact), solves it once, and hands the result to plain Newton; the driver
keeps x_prev instead when the warm start's residual is not smaller.  The
network is trained on the GPU, by the project's own forward / backward /
Adam kernels, on the (features, converged nu) pairs of the first steps of
the same run.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _lib as L
from .material import MaterialData

B_SCALE = 2.2     # T: feature scale of B
S_SCALE = 4.84    # T^2: feature scale of |B|^2
N_IN = 16
NPARAM = 9473
NBF16 = 9216


def _glorot(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in, bias in ((64, 16, True), (64, 64, True), (64, 64, True)):
        lim = math.sqrt(6.0 / (fan_in + fan_out))
        parts += [rng.uniform(-lim, lim, fan_out * fan_in), np.zeros(fan_out)]
    lim = math.sqrt(6.0 / 65.0)
    parts += [rng.uniform(-lim, lim, 64), np.zeros(1)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAM
    return p


class NuPredictor:
    """Per-element reluctivity predictor for the elements of nonlinear regions."""

    def __init__(self, assembler, matLib: dict, seed: int = 0, drive=None):
        self.asm = assembler
        mesh = assembler.Mesh
        self.dev = torch.device("cuda", torch.cuda.current_device())
        nonlin = sorted(int(k) for k, v in matLib.items() if isinstance(v, MaterialData)
                        and v.Type == "Nonlinear")
        if not nonlin:
            raise ValueError("no nonlinear material: nothing to predict")
        self.tags = nonlin
        self.drive = drive
        tags = mesh.RegionTags
        sel = np.isin(tags, nonlin)
        self.idx = torch.from_numpy(np.nonzero(sel)[0].astype(np.int64)).to(self.dev)
        # region code of each selected element: (j + 1) / (number of regions)
        code = np.zeros(int(sel.sum()), dtype=np.float32)
        for j, t in enumerate(nonlin):
            code[tags[sel] == t] = (j + 1.0) / len(nonlin)
        self.region = torch.from_numpy(code).to(self.dev)
        self.tags_sel = torch.from_numpy(tags[sel].astype(np.int32)).to(self.dev)
        self.mat, self._keep = assembler.material_tables(matLib)
        # reluctivity of the linear regions' elements (their constant nu)
        nu0 = np.zeros(int(tags.max()) + 1)
        for k, v in matLib.items():
            if isinstance(v, MaterialData):
                nu0[int(k)] = v.Nu_Linear if v.Type == "Linear" else v.nu0
            else:
                nu0[int(k)] = float(v)
        self.nu_lin = torch.from_numpy(nu0[tags]).to(self.dev)
        self.params = torch.from_numpy(_glorot(seed)).to(self.dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.wb = torch.empty(NBF16, dtype=torch.int16, device=self.dev)
        L.check(L.lib.mpk_mlp_sync_bf16(L.ptr(self.params), L.ptr(self.wb), L.stream_ptr()))
        self.X, self.T = [], []
        # converged B of the selected elements, most recent first (zeros
        # before the first steps: the transient starts from A = 0)
        self.hist = [torch.zeros((self.idx.numel(), 3), dtype=torch.float64, device=self.dev)
                     for _ in range(3)]
        self.t_mean, self.t_std = 0.0, 1.0
        self.adam_step = 0
        self.trained = False
        self.train_info = {}

    # ------------------------------------------------------------- features
    def observe(self, B: torch.Tensor) -> None:
        """Records a converged step's B (all elements)."""
        self.hist = [B.index_select(0, self.idx).to(torch.float64)] + self.hist[:2]

    def features(self, t_next: float = 0.0) -> torch.Tensor:
        b1, b2, b3 = self.hist
        be = 2.0 * b1 - b2
        db = b1 - (2.0 * b2 - b3)
        bq = 3.0 * b1 - 3.0 * b2 + b3
        X = torch.zeros((self.idx.numel(), N_IN), dtype=torch.float32, device=self.dev)
        X[:, 0:3] = (b1 / B_SCALE).float()
        X[:, 3:6] = (be / B_SCALE).float()
        X[:, 6:9] = (db * (10.0 / B_SCALE)).float()
        X[:, 9] = ((b1 * b1).sum(dim=1) / S_SCALE).float()
        X[:, 10] = ((be * be).sum(dim=1) / S_SCALE).float()
        if self.drive is not None:
            d0, d1 = self.drive(t_next)
            X[:, 11] = float(d0)
            X[:, 12] = float(d1)
        X[:, 13] = ((db * db).sum(dim=1) * (100.0 / S_SCALE)).float()
        X[:, 14] = ((bq * bq).sum(dim=1) / S_SCALE).float()
        X[:, 15] = self.region
        return X

    def log_nu(self, B: torch.Tensor, selected: bool = False) -> torch.Tensor:
        """log nu(|B|^2) of the nonlinear-region elements from the B-H tables
        (B over all elements, or over those elements when ``selected``)."""
        Bs = B if selected else B.index_select(0, self.idx)
        s = (Bs ** 2).sum(dim=1).contiguous()
        nu = torch.empty_like(s)
        L.check(L.lib.mpk_eval_nu(L.ptr(s), L.ptr(self.tags_sel), s.numel(), C.byref(self.mat),
                                  L.ptr(nu), None, L.stream_ptr()))
        return torch.log(nu)

    def _base(self) -> torch.Tensor:
        """log nu at the extrapolated field 2 B1 - B2."""
        return self.log_nu(2.0 * self.hist[0] - self.hist[1], selected=True)

    def add_sample(self, B_conv: torch.Tensor, t_next: float = 0.0) -> None:
        """Adds the (features, target) pair of a converged step to the training
        set and records the step."""
        self.X.append(self.features(t_next))
        self.T.append((self.log_nu(B_conv) - self._base()).float())
        self.observe(B_conv)

    # ------------------------------------------------------------- training
    def fit(self, iters: int = 20000, batch: int = 8192, lr: float = 5e-3, seed: int = 0,
            lr_final: float = 0.05):
        X = torch.cat(self.X)
        T = torch.cat(self.T)
        self.t_mean = float(T.mean())
        self.t_std = float(T.std()) or 1.0
        Tn = ((T - self.t_mean) / self.t_std).contiguous()
        n = X.shape[0]
        b = min(batch, n)
        ws = C.c_int64()
        L.check(L.lib.mpk_mlp_work_size(b, C.byref(ws)))
        work = torch.empty(ws.value, dtype=torch.float32, device=self.dev)
        g = torch.empty(NPARAM, dtype=torch.float32, device=self.dev)
        loss = torch.empty(1, dtype=torch.float32, device=self.dev)
        gen = torch.Generator(device=self.dev)
        gen.manual_seed(seed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        first = None
        for it in range(iters):
            sel = torch.randint(0, n, (b,), device=self.dev, generator=gen)
            xb = X.index_select(0, sel).contiguous()
            tb = Tn.index_select(0, sel).contiguous()
            L.check(L.lib.mpk_mlp_loss_grad(L.ptr(self.params), L.ptr(self.wb), L.ptr(xb),
                                            L.ptr(tb), b, L.ptr(g), L.ptr(loss), L.ptr(work),
                                            L.stream_ptr()))
            if first is None:
                first = float(loss)
            self.adam_step += 1
            step_lr = lr * (lr_final ** (it / max(iters, 1)))
            L.check(L.lib.mpk_mlp_adam(L.ptr(self.params), L.ptr(self.m), L.ptr(self.v), L.ptr(g),
                                       self.adam_step, step_lr, 0.9, 0.999, 1e-8, L.ptr(self.wb),
                                       L.stream_ptr()))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # full-set loss (normalised units) after training
        yhat = self._forward(X)
        mse = float(((yhat - Tn) ** 2).mean())
        self.trained = True
        self.train_info = dict(samples=n, iters=iters, batch=b, first_loss=first, mse=mse,
                               rms_log_nu=math.sqrt(mse) * self.t_std,
                               train_ms=1e3 * (t1 - t0))
        return self.train_info

    def _forward(self, X: torch.Tensor) -> torch.Tensor:
        y = torch.empty(X.shape[0], dtype=torch.float32, device=self.dev)
        L.check(L.lib.mpk_mlp_forward(L.ptr(self.params), L.ptr(self.wb), L.ptr(X.contiguous()),
                                      X.shape[0], L.ptr(y), L.stream_ptr()))
        return y

    def predict_nu(self, t_next: float = 0.0) -> torch.Tensor:
        """nu_hat for every element at the next step: the prediction in
        nonlinear regions, nu0 elsewhere."""
        y = self._forward(self.features(t_next))
        nu = self.nu_lin.clone()
        nu[self.idx] = torch.exp(self._base() + y.double() * self.t_std + self.t_mean)
        return nu


def learned_warm_start(predictor: NuPredictor, space, eta_floor: float = 0.5,
                       eta_warm: float = 1e-2):
    """warm_start hook for TransientSolver.solve (None until the network is
    trained).  With nu_hat predicted from the recorded flux history (the
    predictor's on_step hook, training_collector, keeps it), it solves the
    secant system A x = b, A = K(nu_hat) + M_sigma/dt + reg M_reg, for the
    correction around the extrapolated state x_e = 2 x_prev - x_prev2:
    A d = b - A x_e, x0 = x_e + d, to relative accuracy eta_warm (never
    tighter than the step's Newton target, eta_floor * max(Tolerance,
    RelTolerance |R(x_prev)|) absolute): Newton polishes what the prediction
    and the solve leave.  Measured on a 240k-tet E-core transient, warm
    tolerances of 3e-3 and 3e-2 take as many PCG iterations per step as 1e-2
    or more."""

    def hook(ctx):
        if not predictor.trained:
            return None
        s = ctx["solver"]
        x_prev, x_prev2, b, dt = ctx["x_prev"], ctx["x_prev2"], ctx["b"], ctx["dt"]
        nu = predictor.predict_nu(ctx["t"]).contiguous()
        kv = torch.empty_like(s.jsys)
        L.check(L.lib.mpk_csc_curl_curl(s.plan.handle, None, L.ptr(nu), C.byref(s.qk),
                                        L.PREC_FP64, L.ptr(kv), s.sp))
        s._constant_part(dt)
        s._combine(kv, 1.0, kv, 1.0, s.cvals)
        xe = 2.0 * x_prev - x_prev2 if ctx["step"] > 2 else x_prev
        s._spmv(kv, xe, s.tmp)
        r = (b - s.tmp) * s.free
        rn = float(torch.linalg.vector_norm(r))
        target = max(s.Tolerance, s.RelTolerance * ctx["r_prev"])
        if rn <= eta_floor * target:
            return xe
        eta = min(max(eta_floor * target / rn, eta_warm), 1e-1)
        d, nit = s.solve_linear_system(kv, r, eta)
        ctx["warm_pcg"] = nit
        return xe + d

    return hook


def training_collector(predictor: NuPredictor, space, train_steps: int, train_iters: int = 20000,
                       log=None):
    """on_step hook for TransientSolver.solve: records every converged step
    in the predictor's history; over the first ``train_steps`` steps it adds
    the (features, converged log nu) pairs to the training set, and after
    the last of them trains the network (the warm-start hook stays inactive
    until then)."""
    asm = predictor.asm

    def on_step(ctx):
        k = ctx["step"]
        bx = asm.computeElementB(ctx["x"], space)
        if k > train_steps:
            predictor.observe(bx)
            return
        predictor.add_sample(bx, ctx["t"])
        if k == train_steps:
            info = predictor.fit(iters=train_iters)
            if log is not None:
                log(f"[learned] trained: {info}")

    return on_step
