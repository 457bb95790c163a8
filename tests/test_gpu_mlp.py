"""MFMA MLP kernels against the CPU oracle (BF16 rounding points reproduced)."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _setup(p):
    import torch
    from magpackage_amd import _lib as L
    pd = torch.from_numpy(p.astype(np.float32)).cuda()
    wb = torch.empty(9216, dtype=torch.int16, device="cuda")
    L.check(L.lib.mpk_mlp_sync_bf16(L.ptr(pd), L.ptr(wb), L.stream_ptr()))
    return pd, wb


def _round32(p):
    return p.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n", [1, 16, 45, 4096, 70001])
def test_forward_matches_oracle(gpu, n):
    import torch
    from magpackage_amd import _lib as L
    rng = np.random.default_rng(n)
    p = _round32(MO.init_params(1) + rng.normal(size=MO.NPARAM) * 0.02)
    X = rng.normal(size=(n, MO.N_IN)).astype(np.float32)
    pd, wb = _setup(p)
    Xd = torch.from_numpy(X).cuda()
    y = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    L.check(L.lib.mpk_mlp_forward(L.ptr(pd), L.ptr(wb), L.ptr(Xd), n, L.ptr(y), L.stream_ptr()))
    yr, _ = MO.forward(p, X.astype(np.float64), bf=True)
    yg = y.cpu().numpy().astype(np.float64)
    assert np.isfinite(yg).all()
    # FP32 sums against the oracle's FP64 ones: agreement to ~1e-7 except for
    # the rare rows where an activation lands on the other side of a BF16
    # rounding boundary (one BF16 ulp, 2^-8 relative, in that activation)
    err = np.abs(yg - yr)
    scale = max(1.0, np.abs(yr).max())
    assert np.quantile(err, 0.99) <= 1e-5 * scale
    assert err.max() <= 1e-2 * scale


@pytest.mark.parametrize("n", [16, 333, 4096, 50000])
def test_loss_grad_matches_oracle(gpu, n):
    import torch
    from magpackage_amd import _lib as L
    rng = np.random.default_rng(100 + n)
    p = _round32(MO.init_params(2) + rng.normal(size=MO.NPARAM) * 0.02)
    X = rng.normal(size=(n, MO.N_IN)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    pd, wb = _setup(p)
    Xd, td = torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda()
    ws = C.c_int64()
    L.check(L.lib.mpk_mlp_work_size(n, C.byref(ws)))
    work = torch.empty(ws.value, dtype=torch.float32, device="cuda")
    g = torch.empty(MO.NPARAM, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    L.check(L.lib.mpk_mlp_loss_grad(L.ptr(pd), L.ptr(wb), L.ptr(Xd), L.ptr(td), n, L.ptr(g),
                                    L.ptr(loss), L.ptr(work), L.stream_ptr()))
    lr, gr = MO.loss_grad(p, X.astype(np.float64), t.astype(np.float64), bf=True)
    gg = g.cpu().numpy().astype(np.float64)
    assert abs(float(loss) - lr) <= 1e-4 * lr
    err = np.abs(gg - gr).max() / np.abs(gr).max()
    print(f"n={n}: loss {float(loss):.6g} vs {lr:.6g}, grad rel err {err:.2e}", flush=True)
    assert err < 5e-3
    # per block of parameters too (the small blocks must not hide in the max)
    for name, sl in (("W1", slice(0, 1024)), ("b1", slice(1024, 1088)), ("W2", slice(1088, 5184)),
                     ("b2", slice(5184, 5248)), ("W3", slice(5248, 9344)),
                     ("b3", slice(9344, 9408)), ("w4", slice(9408, 9472)),
                     ("b4", slice(9472, 9473))):
        e = np.abs(gg[sl] - gr[sl]).max() / max(np.abs(gr[sl]).max(), 1e-30)
        assert e < 2e-2, (name, e)
    # deterministic: a second call is bitwise identical
    g2 = torch.empty_like(g)
    L.check(L.lib.mpk_mlp_loss_grad(L.ptr(pd), L.ptr(wb), L.ptr(Xd), L.ptr(td), n, L.ptr(g2),
                                    L.ptr(loss), L.ptr(work), L.stream_ptr()))
    assert torch.equal(g.view(torch.int32), g2.view(torch.int32))


def test_adam_training_reduces_loss(gpu):
    """Fit log nu(|B|^2) of the Marrocco curve from (B, |B|^2) features."""
    import torch
    from magpackage_amd import _lib as L
    rng = np.random.default_rng(7)
    n = 8192
    s = rng.uniform(0.0, 4.84, n)
    X = np.zeros((n, MO.N_IN), dtype=np.float32)
    X[:, 0] = np.sqrt(s) / 2.2
    X[:, 1] = s / 4.84
    X[:, 2] = 1.0
    logt = np.log(5e-4 * np.exp(5.0 * s) + 2e-4)
    t = ((logt - logt.mean()) / logt.std()).astype(np.float32)
    p = MO.init_params(9).astype(np.float32)
    pd, wb = _setup(p.astype(np.float64))
    m = torch.zeros_like(pd)
    v = torch.zeros_like(pd)
    Xd, td = torch.from_numpy(X).cuda(), torch.from_numpy(t).cuda()
    ws = C.c_int64()
    L.check(L.lib.mpk_mlp_work_size(n, C.byref(ws)))
    work = torch.empty(ws.value, dtype=torch.float32, device="cuda")
    g = torch.empty(MO.NPARAM, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    losses = []
    for step in range(1, 401):
        L.check(L.lib.mpk_mlp_loss_grad(L.ptr(pd), L.ptr(wb), L.ptr(Xd), L.ptr(td), n, L.ptr(g),
                                        L.ptr(loss), L.ptr(work), L.stream_ptr()))
        if step in (1, 400):
            losses.append(float(loss))
        L.check(L.lib.mpk_mlp_adam(L.ptr(pd), L.ptr(m), L.ptr(v), L.ptr(g), step, 3e-3, 0.9,
                                   0.999, 1e-8, L.ptr(wb), L.stream_ptr()))
    print("loss", losses, flush=True)
    assert losses[1] < 0.01 * losses[0]
