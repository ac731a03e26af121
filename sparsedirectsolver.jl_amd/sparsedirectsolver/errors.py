"""Exception types of the reference (src/errors.jl:10-57)."""


class InvalidValueError(Exception):
    """An invalid argument or parameter value (cuDSS CUDSS_STATUS_INVALID_VALUE)."""


class NotSupportedError(Exception):
    """A feature or combination that is not supported (CUDSS_STATUS_NOT_SUPPORTED)."""


class FactorizationError(Exception):
    """A phase run out of order, or a failed phase (info = the failed column, 0 if none)."""

    def __init__(self, info, msg):
        super().__init__(msg)
        self.info = info


class InterruptedError_(Exception):
    """The user_host_interrupt flag was set while a phase ran (src/logging.jl:59-63)."""
