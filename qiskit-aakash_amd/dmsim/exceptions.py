"""Exception types (the reference's BasicAerError, qiskit/providers/basicaer/exceptions.py)."""


class BasicAerError(Exception):
    """Raised for invalid options, circuits or states, as in the reference."""

    def __init__(self, *message):
        super().__init__(' '.join(str(m) for m in message))
        self.message = ' '.join(str(m) for m in message)
