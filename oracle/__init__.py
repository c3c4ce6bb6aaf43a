"""CPU oracle: a numpy restatement of the reference's CPU evaluation path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports, calls or
links anything under ``oracle/``; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and there only as the
checker (or as the timed CPU baseline), never as the measured GPU path.

Each function cites the reference file:line it restates (paths relative to the
reference checkout).  The reference's CPU path is PyTorch CPU ops; this
restatement uses numpy with the same operation order wherever the order fixes a
floating point result:

* ``iou_sum`` accumulation is an in-order left fold (``np.add.at``), matching
  the reference's 1-D ``scatter_add_`` (evaluator.py:413-415).
* Every ratio is one IEEE float64 division of the same operands.
* Group means (All/Things/Stuff, mIoU, AP/AR summaries) are sums whose order
  the reference leaves to torch's reduction; they agree to ~1e-16 relative,
  well inside the 1e-12 relative bar.

Parity pins (see tests/test_oracle_*.py): the reference's own known-answer
tests restated, brute-force checks on tiny inputs, invariants, and golden
vectors written from the reference in the development container
(tests/golden/make_golden.py).
"""

from . import panoptic, instance, combined, prefix, rle  # noqa: F401
