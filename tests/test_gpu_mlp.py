"""Learned null: MLP kernels against the FP64 oracle, training, end to end."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tenex_amd import core, synthetic
from tenex_amd.inference.learned_null import NullMLP

pytestmark = pytest.mark.gpu


def _params_as_used(model):
    """Master parameters with weights rounded to the bf16 the kernels read."""
    w1, b1, w2, b2, w3, b3 = [v.cpu().double().numpy() for v in model.views()]
    to = lambda t: (t.view(torch.int32) & -65536).view(torch.float32) if False else None  # noqa
    def bf(x):
        u = torch.from_numpy(x.astype(np.float32)).view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.to(torch.int32).view(torch.float32).double().numpy()
    return np.concatenate([bf(w1), b1, bf(w2), b2, bf(w3), b3])


@pytest.mark.parametrize("N", [1, 300, 20000])
def test_mlp_forward_and_gradient_vs_oracle(dev, N):
    """N = 20000 spans more 64-row chunks than workgroups, so workgroups
    accumulate several chunks before their partial rows are summed."""
    D, Oo = 17, 2
    m = NullMLP(D, Oo, dev, seed=3)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((N, D)).astype(np.float32)
    T = rng.standard_normal((N, Oo)).astype(np.float32)
    Xd, Td = torch.from_numpy(X).to(dev), torch.from_numpy(T).to(dev)
    p = _params_as_used(m)
    Y = m.forward(Xd).cpu().numpy()
    Yo = O.mlp_forward(p, D, 64, Oo, X.astype(np.float64))
    assert np.abs(Y - Yo).max() < 1e-4 * max(1.0, np.abs(Yo).max())
    loss, g = m.loss_grad(Xd, Td)
    lo, go = O.mlp_loss_grad(p, D, 64, Oo, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss.item() - lo) < 1e-4 * max(1.0, lo)
    g = g.cpu().numpy()
    scale = np.abs(go).max()
    assert np.abs(g - go).max() < 1e-4 * scale


def test_mlp_gradient_deterministic_and_wide_input(dev):
    D, Oo, N = 64, 4, 5000
    m = NullMLP(D, Oo, dev, seed=7)
    rng = np.random.default_rng(1)
    X = torch.from_numpy(rng.standard_normal((N, D)).astype(np.float32)).to(dev)
    T = torch.from_numpy(rng.standard_normal((N, Oo)).astype(np.float32)).to(dev)
    l1, g1 = m.loss_grad(X, T)
    l2, g2 = m.loss_grad(X, T)
    assert torch.equal(g1, g2) and torch.equal(l1, l2)
    p = _params_as_used(m)
    lo, go = O.mlp_loss_grad(p, D, 64, Oo, X.cpu().double().numpy(), T.cpu().double().numpy())
    assert np.abs(g1.cpu().numpy() - go).max() < 1e-4 * np.abs(go).max()


def test_native_fit_matches_stepwise_schedule(dev):
    """tnx_mlp_fit (host loop in the library) against backward + Adam driven
    step by step from Python with the same cosine schedule; learning rates
    are powers of two so both sides form identical float32 rates."""
    import math
    rng = np.random.default_rng(3)
    X = torch.from_numpy(rng.standard_normal((3000, 9)).astype(np.float32)).to(dev)
    T = torch.from_numpy(rng.standard_normal((3000, 2)).astype(np.float32)).to(dev)
    lr, lr_min, steps = 2.0 ** -9, 2.0 ** -17, 20
    a = NullMLP(9, 2, dev, seed=5)
    la = a.fit(X, T, steps=steps, lr=lr, lr_min=lr_min)
    b = NullMLP(9, 2, dev, seed=5)
    lb = []
    for s in range(steps):
        loss, g = b.loss_grad(X, T)
        lb.append(float(loss.item()))
        rate = lr_min + (lr - lr_min) * 0.5 * (1 + math.cos(math.pi * s / steps))
        b.adam_step(g, lr=rate)
    assert la[0] == lb[0] and la[-1] == lb[-1]
    assert torch.equal(a.master, b.master) and torch.equal(a.W2, b.W2)


def test_native_minibatch_fit_matches_stepwise(dev):
    """Minibatch steps of tnx_mlp_fit: rows shuffled once, step s takes the
    (s mod nb)-th run of `batch` rows (the tail past nb * batch unused);
    the same steps driven from Python give identical weights."""
    import math
    rng = np.random.default_rng(4)
    X = torch.from_numpy(rng.standard_normal((2600, 9)).astype(np.float32)).to(dev)
    T = torch.from_numpy(rng.standard_normal((2600, 2)).astype(np.float32)).to(dev)
    lr, lr_min, steps, batch = 2.0 ** -9, 2.0 ** -17, 7, 1000
    a = NullMLP(9, 2, dev, seed=6)
    la = a.fit(X, T, steps=steps, lr=lr, lr_min=lr_min, batch=batch, seed=9)
    g = torch.Generator(device=dev).manual_seed(9)
    order = torch.randperm(2600, generator=g, device=dev)
    Xs, Ts = X[order].contiguous(), T[order].contiguous()
    b = NullMLP(9, 2, dev, seed=6)
    lb = []
    for s in range(steps):
        r0 = (s % 2) * batch
        loss, gr = b.loss_grad(Xs[r0:r0 + batch].contiguous(), Ts[r0:r0 + batch].contiguous())
        lb.append(float(loss.item()))
        rate = lr_min + (lr - lr_min) * 0.5 * (1 + math.cos(math.pi * s / steps))
        b.adam_step(gr, lr=rate)
    assert la[0] == lb[0] and la[-1] == lb[-1]
    assert torch.equal(a.master, b.master)


def test_mlp_adam_and_bf16_copies(dev):
    m = NullMLP(8, 2, dev, seed=1)
    g = torch.ones(m.n_params, device=dev) * 0.5
    before = m.master.clone()
    m.adam_step(g, lr=1e-2)
    torch.cuda.synchronize()
    assert torch.allclose(m.master, before - 1e-2, atol=1e-6)  # first Adam step = -lr*sign
    w1 = m.views()[0]
    u = w1.view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    exp = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).to(torch.int32).to(torch.int16)
    assert torch.equal(m.W1, exp)


def test_mlp_fit_learns_smooth_target(dev):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((6000, 6)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1] ** 2, X[:, 2] * X[:, 3]], 1).astype(np.float32)
    m = NullMLP(6, 2, dev, seed=0)
    losses = m.fit(torch.from_numpy(X).to(dev), torch.from_numpy(T).to(dev), steps=600, lr=3e-3)
    assert losses[-1] < 0.1 * losses[0]


def test_learned_surrogate_end_to_end(dev):
    import tenex_amd as tnx
    n, T = 120, 1200
    x = synthetic.zero_inflated_lognormal(n, T, seed=12)
    rng = np.random.default_rng(5)
    for s, t in [(1, 2), (10, 11), (20, 33), (40, 41), (50, 60)]:
        x[t, 1:] = x[s, :-1] + 0.2 * rng.random(T - 1).astype(np.float32)
    names = np.array([f"g{i}" for i in range(n)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="gpu")
    nw = tnx.NetWeaver(res, fdr=0.05)
    plain = nw.infer("surrogate_test", n_surrogates=30)
    lr = nw.infer("surrogate_test", n_surrogates=30, learned="mlp",
                  learned_kwargs=dict(train_frac=0.1, min_train=1000, steps=300))
    info = lr.learned_info
    assert info["plain_fraction"] < 0.6
    a = set(map(tuple, plain.grn.pairs.tolist()))
    b = set(map(tuple, lr.grn.pairs.tolist()))
    jac = len(a & b) / max(1, len(a | b))
    assert jac >= 0.9, (jac, len(a), len(b))


def test_learned_host_and_device_pair_lists_agree(dev, monkeypatch):
    """Sources and targets given: the learned null builds its pair lists on
    the host; forcing the device path (nonzero over the mask) gives the same
    p-values, edges and re-run set."""
    import tenex_amd as tnx
    from tenex_amd.inference import learned_null as LN
    n, T = 150, 1000
    x = synthetic.zero_inflated_lognormal(n, T, seed=14)
    rng = np.random.default_rng(6)
    for s, t in [(1, 2), (10, 11), (20, 33), (40, 41)]:
        x[t, 1:] = x[s, :-1] + 0.2 * rng.random(T - 1).astype(np.float32)
    names = np.array([f"g{i}" for i in range(n)])
    src = names[[1, 10, 20, 40, 55, 70, 90, 100, 120, 140, 2]]
    tgt = names[np.r_[2:60, 100:150]]              # overlaps the sources
    res = tnx.TransferEntropyEngine(x, names, sources=src).compute(accelerator="gpu")
    nw = tnx.NetWeaver(res, fdr=0.05)
    kw = dict(n_surrogates=20, targets=tgt, pairs_scope="sources", learned="mlp",
              learned_kwargs=dict(min_train=300, steps=100))
    a = nw.infer("surrogate_test", **kw)
    orig = LN.learned_surrogate_test
    monkeypatch.setattr(LN, "learned_surrogate_test",
                        lambda *args, **k: orig(*args, **{**k, "tested": None}))
    b = nw.infer("surrogate_test", **kw)
    assert a.learned_info["n_rerun"] == b.learned_info["n_rerun"]
    assert np.array_equal(a.p_values, b.p_values)
    assert np.array_equal(a.grn.pairs, b.grn.pairs)
    assert len(a.grn.pairs) > 0
