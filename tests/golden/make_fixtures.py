"""Write the reference-test fixtures used by the parity tests as .npz data files.

Reads the numeric literals (and only those) out of the reference's pytest files with
`ast.literal_eval` on the argument nodes; nothing from the reference is imported or run.
Run once in a container that has the reference checked out at /root/reference:

    python tests/golden/make_fixtures.py

Outputs (committed):
  dg24.npz  the 24-unknown SIPG P1 2D system of src/dd_solvers/tests/test_solvers.py:25-108
            (CSR crow/col/values, load vector, fine_to_solvers, solvers_to_coarse)
  bsr8.npz  the 8-block-row BSR matrix (3x3 blocks) of src/dd_solvers/tests/test_bsr.py:9-138
"""

from __future__ import annotations

import ast
import os

import numpy as np

REF = "/root/reference/src/dd_solvers/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def _kwargs_of_call(node: ast.Call) -> dict:
    return {kw.arg: kw.value for kw in node.keywords}


def _first_arg_literal(call: ast.AST):
    """Literal list inside torch.tensor([...], ...) / torch.Size((...))."""
    assert isinstance(call, ast.Call), ast.dump(call)
    return ast.literal_eval(call.args[0])


def _module_assign(tree: ast.Module, name: str) -> ast.AST:
    for node in tree.body:
        if isinstance(node, ast.Assign) and any(
                isinstance(t, ast.Name) and t.id == name for t in node.targets):
            return node.value
    raise KeyError(name)


def dg24():
    with open(os.path.join(REF, "test_solvers.py")) as f:
        tree = ast.parse(f.read())
    call = _module_assign(tree, "form_matrix")
    kw = _kwargs_of_call(call)
    values = np.array(_first_arg_literal(kw["values"]), dtype=np.float64)
    cols = np.array(_first_arg_literal(kw["col_indices"]), dtype=np.int32)
    crow = np.array(_first_arg_literal(kw["crow_indices"]), dtype=np.int32)
    load = np.array(_first_arg_literal(_module_assign(tree, "load_vector")), dtype=np.float64)
    f2s = np.array(_first_arg_literal(_module_assign(tree, "fine_to_solvers")), dtype=np.int32)
    s2c = np.array(_first_arg_literal(_module_assign(tree, "solvers_to_coarse")), dtype=np.int32)
    np.savez(os.path.join(HERE, "dg24.npz"), values=values, col_indices=cols,
             crow_indices=crow, load_vector=load, fine_to_solvers=f2s,
             solvers_to_coarse=s2c, dofs_per_solver=np.array(6),
             fine_to_dofs=np.arange(24, dtype=np.int32).reshape(8, 3))


def bsr8():
    with open(os.path.join(REF, "test_bsr.py")) as f:
        tree = ast.parse(f.read())
    call = _module_assign(tree, "matrix")
    kw = _kwargs_of_call(call)
    crow = np.array(_first_arg_literal(kw["crow_indices"]), dtype=np.int64)
    cols = np.array(_first_arg_literal(kw["col_indices"]), dtype=np.int64)
    values = np.array(_first_arg_literal(kw["values"]), dtype=np.float64)
    np.savez(os.path.join(HERE, "bsr8.npz"), crow_indices=crow, col_indices=cols, values=values)


if __name__ == "__main__":
    dg24()
    bsr8()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
