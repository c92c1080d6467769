"""Tracing storage: the context that gate constructors record themselves into.

Same contract as the reference's ``StorageBase``/``CircuitStorage``
(tedq/QInterpreter/circuits/storage_base.py:28-162): entering the context
pushes it on the process-wide deque, operators created while it is active
append themselves to ``storage_context`` in program order, and measurements
replace the observable they wrap (``remove`` then ``append``).
"""

import abc

from .global_variables import GlobalVariables


class StorageBase(abc.ABC):
    """Abstract context manager that records operators while active."""

    _active_warehouse = GlobalVariables.get_value("global_deque")

    def __enter__(self):
        StorageBase._active_warehouse.append(self)
        return self

    def __exit__(self, exception_type, exception_value, traceback):
        StorageBase._active_warehouse.pop()

    @abc.abstractmethod
    def _append(self, obj, **kwargs):
        """Record ``obj`` at the end of this storage."""

    @abc.abstractmethod
    def _remove(self, obj):
        """Forget ``obj``."""

    @classmethod
    def recording(cls):
        return bool(cls._active_warehouse)

    @classmethod
    def active_context(cls):
        if cls.recording():
            return cls._active_warehouse[-1]
        return None

    @classmethod
    def append(cls, obj, **kwargs):
        if cls.recording():
            cls.active_context()._append(obj, **kwargs)  # pylint: disable=protected-access

    @classmethod
    def remove(cls, obj):
        if cls.recording():
            cls.active_context()._remove(obj)  # pylint: disable=protected-access

    @classmethod
    def update_info(cls, obj, **kwargs):
        if cls.recording():
            cls.active_context()._update_info(obj, **kwargs)  # pylint: disable=protected-access

    def _update_info(self, obj, **kwargs):
        raise NotImplementedError

    @classmethod
    def get_info(cls, obj):
        if cls.recording():
            return cls.active_context()._get_info(obj)  # pylint: disable=protected-access
        return None

    def _get_info(self, obj):
        raise NotImplementedError


class CircuitStorage(StorageBase):
    """Ordered list of everything recorded while the context is active."""

    def __init__(self):
        self.storage_context = []

    def _append(self, obj, **kwargs):
        self.storage_context.append(obj)

    def _remove(self, obj):
        self.storage_context.remove(obj)
