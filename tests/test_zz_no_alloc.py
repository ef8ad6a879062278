"""dds_local_apply_sym_mfma_noalloc (its workspace in the library's static device scratch)
against dds_local_apply_sym_mfma (a workspace the caller allocates): the same kernels, so the
same bits, on two calls in a row with the second on a smaller input; a workspace beyond the
static scratch is refused before any launch."""

import numpy as np
import pytest
import torch

import dd_solvers as dds
from dd_solvers import _lib
from dd_solvers import problems as P

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


def _setup(monkeypatch, N, box, overlap, storage, direct):
    pb = P.make_problem("q1_3d", N, box=box, overlap=overlap, seed=5, lognormal=True)
    sizes = np.diff(pb.decomposition.sub_ptr)
    P_max = -(-(-(-int(sizes.max()) // 64)) // 2)
    monkeypatch.setattr(dds.solvers, "_symm_group", lambda sizes, n_cu: P_max if direct else 1)
    A = torch.sparse_csr_tensor(torch.as_tensor(pb.A.indptr), torch.as_tensor(pb.A.indices),
                                torch.as_tensor(pb.A.data), pb.A.shape).to(dev)
    M = dds.AdditiveSchwarz(torch.float64, dds.Inv(storage, scaled=storage == torch.float16,
                                                   accumulate=torch.float32, layout="sym_mfma"),
                            dds.DenseCholesky())
    M.setup(A, decomposition=pb.decomposition)
    assert M.local.layout == "sym_mfma" and M.local.direct == direct
    x = torch.as_tensor(np.random.default_rng(N).standard_normal(pb.A.shape[0]), device=dev)
    # the entry points write the vectors' dtype (FP64), whatever the dtype the preconditioner
    # keeps its own local results in
    Y = torch.empty(M.Y.shape[0] if M.Y is not None else x.shape[0], dtype=torch.float64, device=dev)
    return M, x, Y


def _noalloc(M, x, Y, n_blk=None):
    L, xp = M.local, M._xpack
    out, pos = (Y, M.sub_ptr) if L.direct else (M.Yp, L.ypos)
    _lib.call("dds_local_apply_sym_mfma_noalloc", _lib.dtype_code(L.mat_dtype),
              _lib.dtype_code(M.pi), M.n_solvers, L.n_items, L.max_k, L.group,
              L.items.data_ptr(), M.sub_ptr.data_ptr(), _lib.ptr(M.sub_idx), L.mat_off.data_ptr(),
              pos.data_ptr(), _lib.ptr(L.scale), L.mats.data_ptr(), x.data_ptr(),
              xp["n_blk"] if n_blk is None else n_blk, xp["blk_owner"].data_ptr(),
              xp["blk_ptr"].data_ptr(), out.data_ptr(), None, _lib.stream())
    if not L.direct:
        M._enqueue_sym_combine(Y, None)


def _with_workspace(M, x, Y):
    M._enqueue_sym_apply(x, None, Y)
    if not M.local.direct:
        M._enqueue_sym_combine(Y, None)


@pytest.mark.parametrize("order", ["large_direct_then_small", "large_groups_then_small_direct"])
def test_noalloc_matches_workspace_entry_point(monkeypatch, order):
    cases = ([(20, 8, 1, torch.float16, True), (14, 4, 1, torch.bfloat16, False)]
             if order == "large_direct_then_small" else
             [(20, 8, 1, torch.bfloat16, False), (10, 4, 1, torch.float16, True)])
    for N, box, overlap, storage, direct in cases:
        print(f"no-alloc apply: N {N} box {box} overlap {overlap} {storage} direct {direct}",
              flush=True)
        M, x, Y1 = _setup(monkeypatch, N, box, overlap, storage, direct)
        Y1.fill_(float("nan"))
        _with_workspace(M, x, Y1)
        Y2 = torch.full_like(Y1, float("nan"))
        _noalloc(M, x, Y2)
        torch.cuda.synchronize()
        assert not torch.isnan(Y1).any()
        assert torch.equal(Y1, Y2)
        M.destroy()


def test_noalloc_refuses_beyond_scratch(monkeypatch):
    M, x, Y = _setup(monkeypatch, 10, 4, 0, torch.bfloat16, True)
    with pytest.raises(_lib.HipError, match="static scratch"):
        _noalloc(M, x, Y, n_blk=(32 << 20) // 388 + 1)
