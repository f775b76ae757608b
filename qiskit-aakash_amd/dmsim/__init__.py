"""dmsim -- MI355X-native engine for the dm_simulator density-matrix path.

The 4^n real Pauli coefficients of an n-qubit density matrix live in HBM and
hand-written CDNA4 HIP kernels (libdmsim.so, C ABI in include/dmsim.h) update
them in place.  The Python host mirrors the reference backend's interface:
``BasicAer.get_backend('dm_simulator')``, ``execute()``, its
``backend_options`` and its result dictionary.
"""
__version__ = '0.1.0'

from .circuit import ClassicalRegister, QuantumCircuit, QuantumRegister  # noqa: E402
from .exceptions import BasicAerError  # noqa: E402
from .providers import BasicAer, execute  # noqa: E402
from .transpiler import assemble, transpile  # noqa: E402

__all__ = ['BasicAer', 'BasicAerError', 'ClassicalRegister', 'QuantumCircuit', 'QuantumRegister',
           'assemble', 'execute', 'transpile']
