"""The setup helpers on device tensors, with the reference's own known answers:
sort_csr and gather_csr (src/dd_solvers/tests/test_cusparse.py:8-78), csr_to_bsr and its
block layout (src/dd_solvers/tests/test_bsr.py), permute_csr_tensor on the DG-24 fixture."""

import numpy as np
import pytest
import torch

from helpers import dg24

from dd_solvers.csr import asm_permutation, csr_to_bsr, gather_csr, inverse_permutation, \
    permute_csr_tensor, sort_csr

pytestmark = pytest.mark.gpu
dev = torch.device("cuda")


@pytest.mark.parametrize("ind_dtype", [torch.int32, torch.int64])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("num_parts", [1, 2, 3, 4, None])
def test_sort_csr(ind_dtype, dtype, num_parts):
    # 1 2 0 0 / 0 3 0 0 / 4 0 5 0 / 6 7 8 9, columns shuffled inside the rows
    M = torch.sparse_csr_tensor(torch.tensor([0, 2, 3, 5, 9], dtype=ind_dtype),
                                torch.tensor([1, 0, 1, 0, 2, 2, 3, 1, 0], dtype=ind_dtype),
                                torch.tensor([2, 1, 3, 4, 5, 8, 9, 7, 6], dtype=dtype)).to(dev)
    S = sort_csr(M, num_parts=num_parts)
    assert S.values().is_cuda
    assert torch.equal(S.crow_indices(), torch.tensor([0, 2, 3, 5, 9], dtype=torch.int32, device=dev))
    assert torch.equal(S.col_indices(), torch.tensor([0, 1, 1, 0, 2, 0, 1, 2, 3], dtype=torch.int32, device=dev))
    assert torch.equal(S.values(), torch.tensor([1, 2, 3, 4, 5, 6, 7, 8, 9], dtype=dtype, device=dev))


def test_gather_csr_simple():
    M = torch.sparse_csr_tensor(torch.tensor([0, 3, 6], dtype=torch.int32),
                                torch.tensor([0, 2, 2, 0, 0, 1], dtype=torch.int32),
                                torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0, 6.0], dtype=torch.float32)).to(dev)
    G = gather_csr(M)
    assert torch.equal(G.crow_indices(), torch.tensor([0, 2, 4], dtype=torch.int32, device=dev))
    assert torch.equal(G.col_indices(), torch.tensor([0, 2, 0, 1], dtype=torch.int32, device=dev))
    assert torch.equal(G.values(), torch.tensor([1.0, 5.0, 9.0, 6.0], dtype=torch.float32, device=dev))


def test_gather_csr_empty_row():
    M = torch.sparse_csr_tensor(torch.tensor([0, 2, 2, 5], dtype=torch.int32),
                                torch.tensor([0, 1, 0, 0, 1], dtype=torch.int32),
                                torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0], dtype=torch.float32)).to(dev)
    G = gather_csr(M)
    assert torch.equal(G.crow_indices(), torch.tensor([0, 2, 2, 4], dtype=torch.int32, device=dev))
    assert torch.equal(G.col_indices(), torch.tensor([0, 1, 0, 1], dtype=torch.int32, device=dev))
    assert torch.equal(G.values(), torch.tensor([1.0, 2.0, 7.0, 5.0], dtype=torch.float32, device=dev))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_csr_to_bsr_known_answer(dtype):
    A = torch.sparse_csr_tensor(torch.tensor([0, 2, 3, 5, 9]), torch.tensor([0, 1, 1, 0, 2, 0, 1, 2, 3]),
                                torch.tensor([1, 2, 3, 4, 5, 6, 7, 8, 9], dtype=dtype)).to(dev)
    B = csr_to_bsr(A, 2)
    assert B.layout == torch.sparse_bsr and B.dtype == dtype and B.values().is_cuda
    assert B.crow_indices().tolist() == [0, 1, 3]
    assert B.col_indices().tolist() == [0, 0, 1]
    assert B.values().tolist() == [[[1, 2], [0, 3]], [[4, 0], [6, 7]], [[5, 0], [8, 9]]]
    assert B.values().stride() == (4, 1, 2)


def test_permute_dg24():
    A, b, f2s, s2c, k = dg24()
    f2d = torch.arange(24).reshape(8, 3)
    perm = asm_permutation(torch.as_tensor(f2s), torch.as_tensor(s2c), f2d)
    inv = inverse_permutation(perm)
    At = torch.sparse_csr_tensor(torch.as_tensor(A.indptr), torch.as_tensor(A.indices),
                                 torch.as_tensor(A.data), A.shape).to(dev)
    Ap = permute_csr_tensor(At, inv.to(dev))
    p = perm.numpy()
    assert np.allclose(Ap.to_dense().cpu().numpy(), A.toarray()[p][:, p])
