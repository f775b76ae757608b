"""Provider and ``execute`` (basicaer/__init__.py:26, basicaerprovider.py:34-125,
execute.py:24-221).

``BasicAer.get_backend('dm_simulator')`` returns one backend instance per
process (the reference's provider caches its backends, so options a run sets
on the backend persist, dm_simulator.py:200-266).
"""
from . import transpiler
from .backend import DmSimulator
from .exceptions import BasicAerError

_ALIASES = {'dm_simulator': 'dm_simulator', 'dm_simulator_py': 'dm_simulator'}


class BasicAerProvider:
    def __init__(self):
        self._backends = {'dm_simulator': DmSimulator(provider=self)}

    def get_backend(self, name=None, **kwargs):
        if name is None:
            return self._backends['dm_simulator']
        resolved = _ALIASES.get(name)
        if resolved is None:
            raise BasicAerError("The '%s' backend is not installed in your system." % name)
        return self._backends[resolved]

    def backends(self, name=None, **kwargs):
        if name is None:
            return list(self._backends.values())
        return [self.get_backend(name)]

    def __str__(self):
        return 'BasicAer'


BasicAer = BasicAerProvider()


def execute(experiments, backend, basis_gates=None, coupling_map=None, backend_properties=None,
            initial_layout=None, seed_transpiler=None, optimization_level=None, pass_manager=None,
            qobj_id=None, qobj_header=None, shots=1, memory=False, max_credits=None, seed_simulator=None,
            default_qubit_los=None, default_meas_los=None, schedule_los=None, meas_level=None,
            meas_return=None, memory_slots=None, memory_slot_size=None, rep_time=None,
            parameter_binds=None, **run_config):
    """Transpile + assemble the circuits, then ``backend.run(qobj,
    backend_options=run_config)`` (execute.py:188-221).  Keyword arguments the
    reference forwards to the transpiler and assembler are accepted; on this
    path only the simulator options in ``run_config`` have an effect."""
    qobj = transpiler.assemble(experiments, qobj_id=qobj_id)
    return backend.run(qobj, backend_options=dict(run_config))
