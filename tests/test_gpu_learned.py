"""What the learned start gains (trained net shipped with the package,
nutrlan_amd/data/trlnet_anderson2d.npy, trained on L = 64, 96, 128, 192,
256).

On 2D Anderson lattices from the batched workload's distribution (L ~
U{64..256}, W ~ U[1, 12], potentials never seen in training), the batched
mixed-precision solver runs twice on the same operators: from the net's
start vectors and from random starts.  Both must find the same 8 lowest
eigenpairs at residual <= 1e-8 ||A||; the learned start must need fewer
operator applications in total and on most lattices.  Measured on MI355X
on the 1024 workload lattices: 3.5 % fewer in total (1.6 % for L < 112,
4.0-4.5 % above)."""
import numpy as np
import pytest

from nutrlan_amd import learned as LN
from nutrlan_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_learned_start_saves_matvecs(gpu_lib):
    T = gpu_lib
    net = LN.load_net()
    assert net is not None, "shipped weights missing"
    rng = np.random.default_rng(2024)
    cnt = 96
    Ls = rng.integers(64, 257, cnt)
    models = [LN.lattice(int(L), float(rng.uniform(1, 12)), 700000 + i) for i, L in enumerate(Ls)]
    ops = [T.Csr(*c, mixed=True) for c, d in models]
    x_rand = [W.start_vector(int(L) ** 2, seed=710000 + i) for i, L in enumerate(Ls)]
    x_net = LN.start_vectors(net, [d for c, d in models])
    net.close()
    runs = {}
    for name, xs in (("random", x_rand), ("learned", x_net)):
        probs = [(A, -1, LN.NEV, 32, 1e-8, 1, 100000, LN.NEV, x) for A, x in zip(ops, xs)]
        rc, res, wall = T.solve_batch(probs)
        assert rc == 0
        runs[name] = res
    for A in ops:
        A.close()
    mv = {k: np.array([r["matvec"] for r in v]) for k, v in runs.items()}
    for (c, d), a, b in zip(models, runs["random"], runs["learned"]):
        assert a["nec"] == LN.NEV and b["nec"] == LN.NEV
        assert np.all(b["resid"][:LN.NEV] <= 1e-8 * b["anrm"])
        assert np.abs(a["eval"][:LN.NEV] - b["eval"][:LN.NEV]).max() <= 1e-10 * a["anrm"]
    ratio = mv["learned"].sum() / mv["random"].sum()
    fewer = int((mv["learned"] < mv["random"]).sum())
    print(f"operator applications: random {mv['random'].sum()}, learned {mv['learned'].sum()} "
          f"(x{ratio:.4f}); learned fewer on {fewer}/{cnt} lattices", flush=True)
    assert ratio < 0.99
    assert fewer >= 0.6 * cnt
