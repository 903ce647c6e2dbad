"""Write small golden fixtures from the reference implementation (CPU only).

Run once in a container that holds the reference source tree at
/root/reference; the outputs are committed under tests/golden/ and the tests
read only those .npz files (the reference itself never travels).

Fixtures
--------
te_c0_100x500.npz      engine.compute(accelerator="cpu") on the 100-gene x
                       500-cell zero-inflated log-normal workload (seed 0):
                       input data, bins, n_per_var, the float32 TE matrix of
                       the reference's scatter_add path.
surrogate_40x400.npz   NetWeaver(result).infer("surrogate_test") on the CPU loop
                       path (torch CPU generator), 40 genes x 400 cells, 20
                       block-shuffle surrogates, block length 25, seed 42,
                       FDR 0.05: observed, mean, std, effective TE, p-values,
                       edge list.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def zero_inflated_lognormal(n, T, seed):
    rng = np.random.default_rng(seed)
    x = rng.lognormal(1.0, 1.0, size=(n, T)).astype(np.float32)
    x[rng.random((n, T)) < 0.3] = 0.0
    return x


def main():
    sys.path.insert(0, REF)
    import tenex as tnx  # the reference package

    x = zero_inflated_lognormal(100, 500, 0)
    names = np.array([f"g{i}" for i in range(100)])
    res = tnx.TransferEntropyEngine(x, names).compute(accelerator="cpu")
    np.savez_compressed(os.path.join(HERE, "te_c0_100x500.npz"), data=x,
                        bins=np.asarray(res.bin_arrs)[:, :, 0].astype(np.int8),
                        n_per_var=np.asarray(res.n_per_var, np.int32),
                        matrix=np.asarray(res.matrix, np.float32), b_max=res.b_max,
                        kernel=str(res.kernel))

    x2 = zero_inflated_lognormal(40, 400, 8)
    x2[7, 1:] = x2[3, :-1] + 0.1 * np.random.default_rng(1).random(399).astype(np.float32)
    names2 = np.array([f"g{i}" for i in range(40)])
    r2 = tnx.TransferEntropyEngine(x2, names2).compute(accelerator="cpu")
    st = tnx.NetWeaver(r2, fdr=0.05).infer(method="surrogate_test", device="cpu",
                                            n_surrogates=20, block_length=25, seed=42)
    np.savez_compressed(os.path.join(HERE, "surrogate_40x400.npz"), data=x2,
                        bins=np.asarray(r2.bin_arrs)[:, :, 0].astype(np.int8),
                        n_per_var=np.asarray(r2.n_per_var, np.int32),
                        observed=st.observed_te, mean=st.mean_surrogate_te,
                        std=st.std_surrogate_te, effective=st.effective_te, p=st.p_values,
                        edges=np.asarray(st.grn.pairs, np.int64), n_surrogates=20,
                        block_length=25, seed=42, fdr=0.05)
    print("golden fixtures written")


if __name__ == "__main__":
    main()
