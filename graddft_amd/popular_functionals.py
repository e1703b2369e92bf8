"""Built-in functionals (``grad_dft/popular_functionals.py:375-393``).

* ``LSDA``  Slater exchange with the spin-polarisation interpolation (``lsda_x_e``, :29-50)
* ``B88``   LSDA exchange plus Becke's 1988 gradient correction (``b88_density`` is
  [lsda_x_e, b88_x_e] under a unit coefficient, :52-103, 278-285)
* ``VWN``   VWN5 correlation (``vwn_c_e``, :141-195)
* ``PW92``  Perdew-Wang 1992 correlation (``pw92_c_e``, :105-139)
* ``LYP``   Lee-Yang-Parr correlation, original form (``lyp_c_e``, :197-269)
* ``B3LYP`` 0.8 LSDA + 0.72 B88 + 0.19 VWN + 0.81 LYP on the grid plus 0.2 exact exchange
  (``b3lyp_coefficients``, :340-347).  The exact-exchange term is the molecule's chi-based
  semi-numerical term when the molecule carries chi (built with omegas), otherwise the packed-ERI
  exchange of its initial density (see :func:`graddft_amd.scf.initial_exchange`).
* ``B3LYP_GRID`` the grid part of B3LYP alone (no exact exchange).
"""
from .functional import Functional, LinearFunctional

LSDA = LinearFunctional((1.0, 0.0, 0.0, 0.0), name="LSDA")
B88 = LinearFunctional((1.0, 1.0, 0.0, 0.0), name="B88")
VWN = LinearFunctional((0.0, 0.0, 1.0, 0.0), name="VWN")
PW92 = LinearFunctional((0.0, 0.0, 1.0, 0.0), correlation="pw92", name="PW92")
LYP = LinearFunctional((0.0, 0.0, 0.0, 1.0), name="LYP")
B3LYP = Functional(name="B3LYP", exact_exchange=0.2)
B3LYP_GRID = Functional(name="B3LYP-grid", exact_exchange=0.0)
