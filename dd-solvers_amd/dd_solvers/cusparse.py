"""Import path of the reference's sparse utilities (src/dd_solvers/dd_solvers/cusparse.py),
kept so that `from dd_solvers.cusparse import csr_to_bsr, sort_csr, gather_csr` keeps
working. The implementations live in dd_solvers.csr and use no vendor sparse library."""

from .csr import csr_to_bsr, gather_csr, sort_csr

__all__ = ["sort_csr", "csr_to_bsr", "gather_csr"]
