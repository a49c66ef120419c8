"""CPU oracle for the spinoza pair-amplitude state update.

Test infrastructure only: imported by ``tests/``, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``, always as the checker.  The
product package (``spinoza_amd``) never imports it.

* ``pyoracle``  -- ctypes binding of ``spz_oracle.c``, a C restatement of the
  reference's serial gate arithmetic, reductions, measurement and sampling.
* ``textbook``  -- an independent dense numpy simulator written from the gate
  definitions, used to pin the C restatement on small states.
"""
