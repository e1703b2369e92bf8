"""Built-in functionals (``grad_dft/popular_functionals.py:375-393``).

``B3LYP``: 0.8 LSDA + 0.72 B88 + 0.19 VWN + 0.81 LYP on the grid plus 0.2 exact exchange
(``b3lyp_coefficients``, ``popular_functionals.py:340-347``).  The exact-exchange term is
evaluated from the packed ERI (see :func:`graddft_amd.scf.exact_exchange_energy`).
``B3LYP_GRID`` is the grid part alone (no exact exchange).
"""
from .functional import Functional

B3LYP = Functional(name="B3LYP", exact_exchange=0.2)
B3LYP_GRID = Functional(name="B3LYP-grid", exact_exchange=0.0)
