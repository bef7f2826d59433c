"""The small coupled case whose oracle output is committed as
tests/golden/coupled_small.npz (regenerate with tests/golden/make_oracle_fixtures.py)."""
from cases import make_case, run_oracle


def small_case():
    # non-cubic lattice, stencils crossing the periodic boundary, tumbling on
    return make_case((7, 5, 6), 30, tau_s=0.3, tau_b=0.6, ns=3, nt=10, alpha=0.2, seed=42, drho=1e-2, du=1e-2,
                     P=1e-2, l=0.5)


def coupled_small(O):
    out, _ = run_oracle(O, small_case())
    return out
