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
``mlp``      numpy forward / backward of the fusion cost model's perceptron.

Pinning status (the tests named here are the pins):

* gate/channel updates (U3, CX without TSP error and with it at <cos> = 1,
  memory noise, reset, X/Y/Z/N collapse, per-gate depolarising): the dense
  Kraus oracle, n <= 3 (``tests/test_oracle.py``);
* TSP-error CX with <cos> < 1: the documented U(a) channel averaged exactly
  over two symmetric angle distributions with the same <cos delta>, affinely
  extrapolated to the reference's substitution <cos 2 delta> = 4 <cos delta> - 3
  (``tests/test_oracle_pins.py``, with planted coefficient mistakes);
* Bell measurement: dense Bell projectors on the mirrored pair (qubits
  n-1-q2, n-1-q1) of random mixed states at n = 4, 5, both pair orders
  (``tests/test_oracle_pins.py``), and the notebook value
  (measurement.ipynb cell 10);
* ensemble / partial probabilities: dense diagonals and the explicit
  Kronecker operator the reference builds (``tests/test_oracle.py``);
* Pauli -> matrix conversion: the dense construction and README / notebook
  fixtures;
* merge / partition: the notebook partition tables
  (``tests/golden/notebook_fixtures.json``) and, semantically, the dense
  oracle (a merged U3 equals the product of its gates up to a phase);
* the perceptron (``mlp``): central finite differences
  (``tests/test_oracle_mlp.py``).
"""
