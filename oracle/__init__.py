"""CPU oracle for the leapfrog flow-map hot path.

TEST INFRASTRUCTURE ONLY.  Imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, where it is the checker (FP64) or the
labelled CPU baseline (FP32).  The product package never imports it, and the
product path fails loudly when its HIP library is missing instead of falling
back here.

Parity status per function (what pins it) is listed in ``lfm_oracle``'s
module docstring.
"""
