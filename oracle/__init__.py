"""CPU oracle for the spatial-join path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.  The product
package ``vibespatial_amd`` never imports it.

Two layers:

* ``liboracle.so`` (``oracle.c``): the C restatement of the reference's exact
  predicate path, plus the OpenMP CPU join that is timed as the CPU baseline.
* ``fraction_ref``: tiny pure-Python restatements in exact rationals
  (``fractions.Fraction``) that pin the C code on small inputs.
"""

from .oracle import *  # noqa: F401,F403
from . import fraction_ref  # noqa: F401
