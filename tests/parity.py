"""Iteration-count parity for BiCGSTAB on problems where the count is a
chaotic function of rounding.

The reference's count (bicgstab_solver.h:138-231) is exact arithmetic only in
name: on the non-normal convection-diffusion systems the oracle alone moves by
+-7% when nothing but its dot-product summation order (OpenMP thread count)
changes.  The GPU computes the reference recurrences with rho' = (r*, r)
expanded by linearity as (r*, s) - omega (r*, t), another rounding of the same
quantity.  `oracle_counts` samples the oracle's count over reduction orders
and over both forms of rho'; `count_ok` accepts a GPU count inside that
sample's range widened by its width on each side (+-1 where every sample
agrees).  The GPU's rounding differs in kind as well (fused multiply-adds in
every product, tree-ordered reductions), and a few dozen CPU samples do not
bound the tails: on batch system 0 the oracle alone spans 96-113 over 16
orders and two forms, and GPU runs have given 94-104.  A wrong recurrence
shows up as non-convergence or a count off by a factor, and above all in the
deterministic check beside this one: early iterates agree to 1e-8.
"""
from oracle import binding as O

ORDERS = (1, 2, 3, 4, 5, 6, 7, 8, 12, 16)


def oracle_counts(a_csr, m_csr, b, tol, max_iter, threads=ORDERS):
    counts = []
    try:
        for mode in (0, 1):
            O.set_rho_mode(mode)
            counts += list(O.iteration_spread(a_csr, m_csr, b, tol, max_iter, threads=threads))
    finally:
        O.set_rho_mode(0)
    return counts


def count_ok(it, counts):
    lo, hi = min(counts), max(counts)
    w = max(1, hi - lo)
    return lo - w <= it <= hi + w
