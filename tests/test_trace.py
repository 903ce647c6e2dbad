"""TRACE (OutTE / InTE over a thresholded network): the oracle against the
reference-written fixture and closed forms, then the HIP path
(tnx_trace_entropy) against the oracle: fixed-point entropy sums bit-exact,
OutTE / InTE within 1e-6."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-6


def load(name):
    return dict(np.load(os.path.join(G, name), allow_pickle=False))


def _horner(vals):
    h = np.uint64(0)
    with np.errstate(over="ignore"):
        for v in vals:
            h = h * O.TRACE_P + np.uint64(v)
    return h


def test_oracle_keys_equal_per_job_horner_hashes():
    """The four keys of the 4-way fusion equal the Horner hash of each
    entropy's own column list (the identity the reference's
    tests/test_point_4way.py checks), per sample."""
    rng = np.random.default_rng(42)
    b = rng.integers(0, 28, size=(12, 60))
    for d in (1, 2, 5, 10):
        conn = list(rng.choice(np.arange(1, 12), d, replace=False))
        L = 59
        for mode in (0, 1):
            k = O.trace_keys(b, 0, conn, mode, tau=1)
            for t in (0, 17, 58):
                yf, yp = b[conn, t + 1], b[conn, t]
                xf, xp = b[0, t + 1], b[0, t]
                if mode == 0:
                    exp = (_horner(list(yf) + list(yp)), _horner(yp),
                           _horner(list(yf) + [xp] + list(yp)), _horner([xp] + list(yp)))
                else:
                    exp = (_horner([xf, xp]), _horner([xp]), _horner([xf, xp] + list(yp)),
                           _horner([xp] + list(yp)))
                assert tuple(int(kk[t]) for kk in k) == tuple(int(e) for e in exp), (d, mode, t)
            assert all(len(kk) == L for kk in k)


def test_oracle_single_link_is_pairwise_te_and_states_agree():
    """OutTE over one target and InTE over one source are the pairwise TE
    of the counts-based oracle; with several links the hashed entropies
    equal those of the state tuples (no collisions)."""
    rng = np.random.default_rng(3)
    b = rng.integers(0, 5, size=(10, 400)).astype(np.int32)
    b[4, 1:] = b[2, :-1]                       # a copy: large TE
    npv = (b.max(axis=1) + 1).astype(np.int32)
    for tau in (1, 2):
        for s, t in ((2, 4), (0, 7), (5, 1)):
            te = O.te_pairs(b, npv, np.array([[t, s]]), tau=tau)[0]
            assert abs(O.trace_te(b, s, [t], 0, tau) - te) < 1e-12
            assert abs(O.trace_te(b, t, [s], 1, tau) - te) < 1e-12
    for conn in ([1, 3], [2, 5, 8], [0, 1, 3, 6, 9]):
        for mode in (0, 1):
            assert abs(O.trace_te(b, 4, conn, mode) - O.trace_te_tuples(b, 4, conn, mode)) < 1e-12
    assert O.trace_te(b, 4, [], 0) == 0.0


def test_golden_trace_oracle_and_network():
    """The oracle's network, OutTE and InTE reproduce the reference's CPU
    TRACE on the fixture; the package's host network builder gives the same
    network."""
    from tenex_amd.inference.trace import build_causal_network
    g = load("trace_80x700.npz")
    M, bins = g["matrix"], g["bins"].astype(np.int32)
    outte, inte, net, pairs = O.trace_infer(M, bins, gpu=False)
    assert np.array_equal(pairs, g["network"])
    assert np.abs(outte - g["outte"]).max() < 1e-6 and np.abs(inte - g["inte"]).max() < 1e-6
    net_h = build_causal_network(M, 100, 2.0, 42, None)
    got = np.array([[s, j] for j in range(80) for s in net_h[j]], np.int64).reshape(-1, 2)
    assert np.array_equal(got, g["network"]) and np.array_equal(g["grn_pairs"], g["network"])


def test_trace_argument_errors():
    import tenex_amd as tnx
    from tenex_amd.inference import get_method
    m = get_method("trace")
    M = np.zeros((3, 3), np.float32)
    with pytest.raises(ValueError):
        m.infer(M, np.array(list("abc")), torch.device("cpu"))
    with pytest.raises(TypeError):
        m.infer(M, np.array(list("abc")), torch.device("cpu"), bin_data=np.zeros((3, 9)), fdr=0.1)
    with pytest.raises(ValueError):
        m.infer(M, np.array(list("abc")), torch.device("cpu"), bin_data=np.zeros((3, 9, 2)))
    assert "trace" in tnx.available_methods()


def _jobs(rng, n, n_jobs, dmax):
    ids = rng.choice(n, n_jobs, replace=False).astype(np.int32)
    conn = []
    for v in ids:
        d = int(rng.integers(1, dmax + 1))
        c = np.sort(rng.choice(np.setdiff1d(np.arange(n), [v]), d, replace=False))
        conn.append(c.astype(np.int64))
    return ids, conn


@pytest.mark.gpu
@pytest.mark.parametrize("T,tau,levels", [(200, 1, 10), (1025, 1, 6), (3001, 2, 4),
                                          (10000, 1, 8), (16385, 1, 3), (700, 17, 12)])
def test_trace_entropy_sums_bit_exact(dev, T, tau, levels):
    """tnx_trace_entropy's fixed-point sums equal the oracle's key counts
    summed through the same table, bit for bit, for OutTE and InTE jobs of
    degree 1..12 (a frequent level 0 in half the rows); TE within 1e-6."""
    from tenex_amd import core
    rng = np.random.default_rng(T + tau)
    n = 40
    b = rng.integers(0, levels, size=(n, T)).astype(np.int32)
    b[::2][rng.random((n // 2, T)) < 0.5] = 0
    b[7, tau:] = b[3, :-tau]
    bins, _, _, T2 = core.upload_bins(b, dev)
    L = T - tau
    tbl = core.xlogx_fx_table(L, dev).cpu().numpy()
    for mode in (0, 1):
        ids, conn = _jobs(rng, n, 12, 12)
        te, sums = core.trace_te(bins, T, tau, mode, ids, conn, return_sums=True)
        te, sums = te.cpu().numpy(), sums.cpu().numpy()
        for j, (v, c) in enumerate(zip(ids, conn)):
            keys = O.trace_keys(b, int(v), c, mode, tau)
            exp = [int(tbl[np.unique(k, return_counts=True)[1]].sum()) for k in keys]
            assert [int(x) for x in sums[j]] == exp, (mode, j)
            assert abs(te[j] - O.trace_te(b, int(v), c, mode, tau)) < TOL, (mode, j)


@pytest.mark.gpu
def test_golden_trace_netweaver_gpu(dev):
    """NetWeaver(result).infer("trace") on the GPU against the reference's
    CPU TRACE: the same network and GRN, OutTE / InTE within 1e-6."""
    import tenex_amd as tnx
    g = load("trace_80x700.npz")
    names = np.array([f"g{i}" for i in range(80)])
    res = tnx.TransferEntropyEngine(g["data"], names).compute(accelerator="gpu")
    assert np.array_equal(res.bin_arrs[:, :, 0], g["bins"])
    assert np.abs(res.matrix - g["matrix"]).max() < 1e-6
    tr = tnx.NetWeaver((g["matrix"], names)).infer("trace", bin_data=g["bins"])
    got = np.array([[s, j] for j in range(80) for s in tr.network[j]], np.int64).reshape(-1, 2)
    assert np.array_equal(got, g["network"])
    assert np.array_equal(np.asarray(tr.grn.pairs, np.int64), g["grn_pairs"])
    assert np.array_equal(tr.grn.te, g["grn_te"])
    assert tr.outte.dtype == np.float32
    assert np.abs(tr.outte - g["outte"]).max() < TOL
    assert np.abs(tr.inte - g["inte"]).max() < TOL
    tr2 = tnx.NetWeaver(res).infer("trace")          # bins and tau from the result
    assert np.abs(tr2.outte - tr.outte).max() < TOL
    assert tr.top_drivers(3)[0][0] == names[int(np.argmax(g["outte"]))]
