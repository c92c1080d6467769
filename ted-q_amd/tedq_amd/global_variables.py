"""Process-wide registry used by circuit tracing.

Mirrors the reference's ``GlobalVariables`` (tedq/global_variables.py:4-25): a
class-level dict holding, among others, the ``"global_deque"`` of active
circuit storages that gate constructors append themselves to.
"""

from collections import deque


class GlobalVariables:
    """Class-level key/value store shared by every module of the package."""

    _global_dict = {}

    @classmethod
    def set_value(cls, key, value):
        cls._global_dict[key] = value

    @classmethod
    def get_value(cls, key, default_value=None):
        return cls._global_dict.get(key, default_value)


# The tracing stack is created at import time (tedq/_set_global_variables.py:7).
if GlobalVariables.get_value("global_deque") is None:
    GlobalVariables.set_value("global_deque", deque())
