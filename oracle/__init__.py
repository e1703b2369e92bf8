"""FP64 CPU oracle for the GradDFT XC grid path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy / torch-CPU float64, the reference algorithm of
XanaduAI/GradDFT for the self-consistent exchange-correlation grid path, each function
citing the reference file:line it follows.  It is the checker for the HIP path and the
CPU baseline timed by ``bench.py``; it is never imported by ``graddft_amd`` itself and is
never the thing measured or shipped.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg may import it.

Pinning (what checks each function against something other than itself) is listed in the
module docstrings; a function with no independent check says "parity unpinned".
"""
