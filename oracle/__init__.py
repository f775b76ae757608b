"""CPU oracle for the dm_simulator density-matrix update path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
directory; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only to check GPU results
(or, in ``bench.py``, as the timed CPU baseline), never as the thing measured
or shipped.

Contents
--------
``refsim``   FP64 numpy restatement of the reference backend
             (``qiskit/providers/basicaer/dm_simulator.py`` and
             ``qiskit/providers/basicaer/basicaertools.py``): the same
             reshapes, full-state copies and slice updates, in the same order,
             plus the host scheduling (``single_gate_merge``, ``partition``).
``dense``    Brute-force 2^n x 2^n complex density matrix with Kraus
             channels.  It pins the Pauli-basis restatement by physics.
``ptm``      Pauli transfer matrices of the channels.

Pinning status (each function's own docstring names its pin):

* gate/channel updates (U3, CX with and without TSP error, memory noise,
  reset, X/Y/Z/N collapse, per-gate depolarising): pinned by the dense
  Kraus oracle (``tests/test_oracle_dense.py``) for n <= 5;
* ensemble / partial probabilities: pinned by dense diagonals and the explicit
  Kronecker operator the reference builds, at n <= 8;
* Pauli -> matrix conversion: pinned by the dense construction and README /
  notebook fixtures;
* merge / partition: pinned by the notebook partition tables
  (``tests/golden/notebook_fixtures.json``) and, semantically, by the dense
  oracle (merged U3 equals the product of its gates up to a phase);
* TSP-error CX with <cos> < 1: the moment substitution <cos 2a> = 4<cos a> - 3
  is pinned only by the formula text (basicaertools.py:329-336): parity
  unpinned beyond the <cos> = 1 case and a linearity check;
* Bell measurement: pinned by the notebook value
  (measurement.ipynb cell 10) and by the mirrored-stride rule read from the
  code text (dm_simulator.py:741); parity unpinned for n > 3 beyond that.
"""
