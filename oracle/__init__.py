"""CPU oracle for the MaoMi optical stage -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy/SciPy, the reference CPU algorithm of MaoMi's
optical stage (maomi/optics.py, maomi/numerics.py, maomi/_propagation.py,
maomi/_optical_math.py, maomi/excitation.py, maomi/fixed.py, maomi/pupils.py).  It is the
checker that the HIP path is compared against, and the algorithm the same-run
CPU baseline times.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import it.  The product package
(maomi_amd) never imports, links or executes anything under oracle/.

Every function carries the reference file:line it follows and how it is
pinned (golden vectors written by tests/golden/make_golden.py from the
reference itself, closed forms, invariants, brute force).  Functions without
such a check say "parity unpinned" in their docstring.

Two precisions are provided by a ``dtype`` argument: ``np.complex64`` mirrors
the reference's float32/complex64 arithmetic (the "reference CPU path"), and
``np.complex128`` is the FP64 oracle.
"""

from .optics_np import *  # noqa: F401,F403
from . import surrogate_np  # noqa: F401
