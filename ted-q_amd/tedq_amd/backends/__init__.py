"""Execution backends: only the PyTorch-facing MI355X engine is provided."""

from .compiled_circuit import CompiledCircuit
from .pytorch_backend import PyTorchBackend, TorchExecute
