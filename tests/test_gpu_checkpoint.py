"""GPU checks of checkpoint/resume (reference SRC/trlcore.c:434-460,
775-790, 1395-1412, 1484-1495 write; 1653-1679 read): a solve cut short by
maxmv leaves a checkpoint (the reference's file format, checked byte-wise
against the compiled reference in tests/test_kit.py); a second solve with
guess = 2 resumes from it and converges to the eigenpairs of an
uninterrupted solve."""
import numpy as np
import pytest

from nutrlan_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_checkpoint_resume(gpu_lib, tmp_path):
    T = gpu_lib
    rp, ci, v = W.laplace3d_csr(14, seed=21)
    n = len(rp) - 1
    x0 = W.start_vector(n, seed=22)
    A = T.Csr(rp, ci, v)
    args = (1, 8, 40, 1e-9, 1)
    full = T.solve_csr(A, *args, 40000, 8, x0)
    assert full["stat"] == 0 and full["nec"] == 8
    cut = int(full["matvec"] * 0.6)
    prefix = str(tmp_path / "cp")
    part = T.solve_csr(A, *args, cut, 8, x0, checkpoint=(1, prefix))
    assert part["stat"] == 0 and part["matvec"] >= cut and part["nec"] < 8
    raw = open(prefix + "0", "rb").read()
    nrow, jnd = np.frombuffer(raw[:8], dtype=np.int32)
    assert nrow == n and 0 < jnd <= 40
    assert len(raw) == 8 + 16 * jnd + 8 * n * (jnd + 1)
    # the saved basis: orthonormal Ritz vectors plus the residual direction
    V = np.frombuffer(raw[8 + 16 * jnd:], dtype=np.float64).reshape(jnd + 1, n).T
    assert np.abs(V[:, :jnd].T @ V[:, :jnd] - np.eye(jnd)).max() <= 1e-10
    res = T.solve_csr(A, *args, 40000, 8, np.zeros(n), resume=prefix)
    A.close()
    assert res["stat"] == 0 and res["nec"] == 8
    anrm = full["anrm"]
    assert np.abs(np.sort(res["eval"][:8]) - np.sort(full["eval"][:8])).max() <= 1e-9 * anrm
    f = W.csr_matvec(rp, ci, v)
    for i in range(8):
        x = res["evec"][:, i]
        assert np.linalg.norm(f(x) - res["eval"][i] * x) <= 1.01e-9 * res["anrm"]
    # resuming costs fewer operator applications than starting over
    assert res["matvec"] < full["matvec"]


def test_checkpoint_missing_file(gpu_lib, tmp_path):
    T = gpu_lib
    rp, ci, v = W.laplace3d_csr(8, seed=1)
    A = T.Csr(rp, ci, v)
    n = len(rp) - 1
    info = T.init_info(n, 30, 1, 4, 1e-8, 1, 1000)
    r = T.solve_csr(A, 1, 4, 30, 1e-8, 1, 1000, 4, np.zeros(n), resume=str(tmp_path / "none"))
    A.close()
    assert r["stat"] in (-8, -212)
