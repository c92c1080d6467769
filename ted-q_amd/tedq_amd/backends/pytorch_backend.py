"""``PyTorchBackend``: the drop-in compiled-circuit callable, executed on MI355X.

Same constructor, call convention, error behaviour and autograd surface as the
reference PyTorch backend (tedq/backends/pytorch_backend.py:44-1223):

* ``PyTorchBackend(backend, circuit, use_cotengra=False, use_jdopttn=False,
  tn_mode=False, hyper_opt=None, tn_simplify=True, **kwargs)`` with kwargs
  ``requires_grad`` (True), ``interface`` ("pytorch", else ValueError) and
  ``diff_method`` ("back_prop" | "param_shift" | "finite_diff");
* ``compiled(*params)``: every argument must be a tensor on one device
  (ValueError otherwise); elements are flattened in argument order onto the
  trainable slots in gate order; the result is ``torch.stack`` of the
  measurement results (expval -> float32, probs -> float32 [2]*k,
  state -> complex64 [2]*n) and mixed shapes raise ValueError;
* ``back_prop`` gradients come from a hand-written adjoint kernel behind a
  ``torch.autograd.Function``; ``param_shift`` uses the gates' shift recipes
  (2 terms, 4 for controlled rotations) evaluated as ONE batched launch;
* ``finite_diff`` raises NotImplementedError as in the reference (:214-218)
  and an unknown method raises Exception at call time (:162-165).

Beyond the reference: ``execute_batch(params_2d)`` evaluates B parameter rows
in one set of launches (shape [B, num_measurements, ...]).  The device is held
per instance, not per class.  Every state-vector forward also reduces
<psi|psi> in fp64 on the device (``last_norm``); ``norm_tol=<float>`` makes
each call raise when max |<psi|psi> - 1| exceeds it.  When every measurement
is a Z-string expectation, trailing gates that map Z-strings to Z-strings
(diagonal gates, CNOTs, gates outside the observables' backward light cone)
are folded into the observables and qubits left without gates are dropped;
``fold=False`` keeps the whole circuit.
"""

import torch

from .compiled_circuit import CompiledCircuit
from ..engine.lowering import lower_circuit
from ..engine.sv import SVExecutor
from ..tools import dec_to_bin

tcomplex = torch.complex64


def _exec_device(device):
    if device is not None and device.type == "cuda":
        return device
    if not torch.cuda.is_available():
        raise RuntimeError("tedq_amd executes on a ROCm GPU (MI355X); no GPU is visible to torch")
    return torch.device("cuda", torch.cuda.current_device())


def _flat_param_row(params, device):
    """Concatenate all call arguments (flattened, real part) into one [1, P] row on ``device``."""
    if not params:
        return torch.zeros((1, 0), dtype=torch.float32, device=device)
    use64 = any(p.dtype in (torch.float64, torch.complex128) for p in params)
    dt = torch.float64 if use64 else torch.float32
    parts = [(p.real if p.is_complex() else p).reshape(-1).to(device=device, dtype=dt) for p in params]
    return torch.cat(parts).reshape(1, -1)


class _AdjointFunction(torch.autograd.Function):
    """Forward on the fused SV engine; backward is the adjoint sweep kernel."""

    @staticmethod
    def forward(ctx, executor, rows, psi_sink=None):  # pylint: disable=arguments-differ
        psi = executor.run_forward(rows, need_grad=True)
        outs = executor.measure(psi)
        if psi_sink is not None:
            psi_sink.append(psi)
        ctx.executor = executor
        ctx.save_for_backward(rows)
        ctx.block = False
        if isinstance(psi, torch.Tensor):
            # the adjoint sweep walks psi back in place: the first backward consumes it (no copy), a
            # second backward through the same graph recomputes the forward
            ctx.psi = psi
            if _scalar_outputs(executor):
                # one [B, M] float32 block (one stack and one conversion instead of one per output)
                ctx.block = True
                return _scalar_block(outs)
        else:                               # contraction mode: a run record, recomputed in backward
            ctx.run = psi
        res = []
        for spec, o in zip(executor.specs, outs):
            if spec.kind in ("zexp", "fexp"):
                res.append(o.to(torch.float32))
            elif spec.kind == "state":
                res.append(o.clone())
            else:
                res.append(o)
        return tuple(res)

    @staticmethod
    def backward(ctx, *grads):  # pylint: disable=arguments-differ
        ex = ctx.executor
        if hasattr(ctx, "run"):
            (rows,) = ctx.saved_tensors
            grad = ex.run_backward(rows, ctx.run, ex.seed_adjoint(ctx.run, grads), grad_dtype=rows.dtype)
            return None, grad, None
        (rows,) = ctx.saved_tensors
        psi = ctx.psi if ctx.psi is not None else ex.run_forward(rows)
        ctx.psi = None
        if ctx.block:
            # columns of one float64 block (views, no copy per output)
            grads = list(grads[0].to(torch.float64).unbind(1))
        grad = ex.adjoint(rows, psi, grads, grad_dtype=rows.dtype)
        return None, grad, None


def _scalar_outputs(executor):
    """Every measurement of a state-vector executor is an expectation value (one number per row)."""
    return bool(executor.specs) and all(s.kind in ("zexp", "fexp") for s in executor.specs)


def _scalar_block(outs):
    """Per-measurement [B] expectation values as one [B, M] float32 tensor."""
    return torch.stack(outs, 1).to(torch.float32)


class TorchExecute(torch.autograd.Function):
    """Parameter-shift autograd function (reference name, pytorch_backend.py:1191-1223).

    Forward evaluates the circuit; backward forms the Jacobian from the gates'
    shift recipes with all shifted parameter rows in one batched launch and
    returns ``dy @ jac``.
    """

    @staticmethod
    def forward(ctx, backend, rows):  # pylint: disable=arguments-differ
        ctx.backend = backend
        ctx.save_for_backward(rows)
        with torch.no_grad():
            return backend._run_rows(rows)[0][0][0]

    @staticmethod
    def backward(ctx, dy):  # pylint: disable=arguments-differ
        (rows,) = ctx.saved_tensors
        jac = ctx.backend.jacobian_param_shift(rows)
        g = dy.reshape(1, -1).to(jac.dtype) @ jac
        return None, g.reshape(rows.shape).to(rows.dtype)


class PyTorchBackend(CompiledCircuit):
    """Compiled circuit executing on the MI355X engine with a PyTorch interface."""

    _device = None

    def __init__(self, backend, circuit, use_cotengra=False, use_jdopttn=False, tn_mode=False, hyper_opt=None,
                 tn_simplify=True, **kwargs):
        super().__init__(backend, circuit, use_cotengra=use_cotengra, use_jdopttn=use_jdopttn, tn_mode=tn_mode,
                         hyper_opt=hyper_opt, tn_simplify=tn_simplify)
        self._requires_grad = kwargs.get("requires_grad", True)
        self._interface = kwargs.get("interface", "pytorch")
        self._diff_method = kwargs.get("diff_method", "back_prop")
        self._tile_bits = kwargs.get("tile_bits", 12)
        self._reg_bits = kwargs.get("reg_bits", 4)
        self._tile_bits_bwd = kwargs.get("tile_bits_bwd", None)
        # compile per-pass kernels for this circuit structure at first use (hiprtc, a few seconds)
        self._jit = bool(kwargs.get("jit", False))
        # contraction mode: externally found paths, one per measurement (SSA or opt_einsum-style linear)
        self._tn_paths = kwargs.get("tn_paths", None)
        self._reg_bits_bwd = kwargs.get("reg_bits_bwd", None)
        # contraction mode: kernel form of GEMM-shaped steps ("4m" default, "3m" Gauss, "4m64" small tiles)
        self._tn_gemm_form = kwargs.get("tn_gemm_form", "4m")
        # state-vector mode: check max |<psi|psi> - 1| (fp64 on the device) after every forward
        self._norm_tol = kwargs.get("norm_tol", None)
        # state-vector mode, Z-string expectations only: fold trailing gates into the observables and
        # drop qubits left without gates (exact; False keeps every gate and qubit)
        self._fold = bool(kwargs.get("fold", True))
        if self._interface != "pytorch":
            raise ValueError(f"{self._interface}: pytroch_backend only supports pytorch interface!")
        self._device = None
        self._circuit = circuit
        self._executor = None
        self._tn_executor = None
        self._states_after_measurement = None

    # ------------------------------------------------------------------
    def check_parameters_torch_device(self, params):
        """All arguments must be tensors on one device (pytorch_backend.py:85-121)."""
        the_same = None
        for par in params:
            try:
                index = par.device.index if par.is_cuda else -1
            except AttributeError as error:
                raise ValueError("input parameters must be type of pytorch tensor!!") from error
            if the_same is None:
                the_same = index
            if the_same != index:
                raise ValueError("input parameters are not in the same device!!")
        self._device = params[0].device if params else None

    def update_device(self, device):
        self._device = device

    @property
    def device(self):
        return self._device

    @property
    def states_after_measurement(self):
        if self._states_after_measurement is None:
            raise ValueError("no states after measurement found! please measure probs with qubits.")
        return self._states_after_measurement

    # ------------------------------------------------------------------
    def __call__(self, *params):
        self.check_parameters_torch_device(params)
        if self._requires_grad is False:
            with torch.no_grad():
                return self.execute(*params)
        if self._diff_method == "back_prop":
            return self.execute(*params)
        if self._diff_method == "param_shift":
            dtypes = [p.dtype for p in params]
            if any(d != dtypes[0] for d in dtypes):
                raise ValueError("input parameters for parameter shift method must have the same data type!")
            return self.param_shift_execute(*params)
        if self._diff_method == "finite_diff":
            return self.finite_diff_execute(*params)
        raise Exception(  # pylint: disable=broad-exception-raised
            f"Differentiation method {self._diff_method} is not supported. "
            f"Supported methods include {{back_prop, param_shift, finite_diff }}")

    def finite_diff_execute(self, *params):
        raise NotImplementedError

    def get_trainable_parameters(self):
        raise NotImplementedError

    # ------------------------------------------------------------------
    def executor(self):
        """The compiled device program (built on first use)."""
        if self._executor is None:
            if self._tn_mode or self._use_cotengra or self._use_jdopttn:
                from ..tn.executor import TNExecutor  # pylint: disable=import-outside-toplevel
                self._executor = TNExecutor(self, paths=self._tn_paths, gemm_form=self._tn_gemm_form)
            else:
                n, units, n_params, pool = lower_circuit(self)
                init = None
                if self._init_state:
                    init = self._init_state.matrix
                self._executor = SVExecutor(n, units, n_params, pool, self._measurements, init_state=init,
                                            k_max=self._tile_bits, r=self._reg_bits,
                                            k_bwd=self._tile_bits_bwd, r_bwd=self._reg_bits_bwd, fold=self._fold)
                self._executor.jit = self._jit
                self._executor.norm_tol = self._norm_tol
        return self._executor

    @property
    def last_norm(self):
        """<psi|psi> of every row of the last state-vector forward, reduced in fp64 on the device
        (compensated sums); a device float64 tensor [B], or None before the first call / in
        contraction mode."""
        ex = self._executor
        return getattr(ex, "last_norm", None) if ex is not None else None

    def _flatten_and_update(self, params):
        views = []
        for par in params:
            for element in par.view(-1, 1):
                views.append(element)
        CompiledCircuit.execute(self, *views)
        if not views and self.trainable_count():
            raise ValueError(
                f"Error!!!! number of parameters are not matched!! required {self.trainable_count()} but 0 are given")

    def _stack(self, results, batch_dim):
        try:
            return torch.stack(results, 1 if batch_dim else 0)
        except RuntimeError as err:
            raise ValueError("You can not have multiple measurements with different shapes!!") from err

    def _run_rows(self, rows):
        """No-autograd evaluation of [B, P] rows -> list of per-measurement outputs [B, ...]."""
        ex = self.executor()
        psi = ex.run_forward(rows)
        outs = ex.measure(psi)
        if isinstance(psi, torch.Tensor) and _scalar_outputs(ex):
            return [_scalar_block(outs)], psi
        res = []
        for spec, o in zip(ex.specs, outs):
            res.append(o.to(torch.float32) if spec.kind in ("zexp", "fexp") else o)
        return [self._stack(res, True)], psi

    def execute(self, *params):
        """Evaluate once with the given arguments (reference semantics, result on the argument device)."""
        self._flatten_and_update(params)
        dev = _exec_device(self._device)
        rows = _flat_param_row(params, dev)
        ex = self.executor()
        sink = []
        if torch.is_grad_enabled() and rows.requires_grad:
            outs = _AdjointFunction.apply(ex, rows, sink)
        else:
            with torch.no_grad():
                psi = ex.run_forward(rows)
                sink.append(psi)
                outs = ex.measure(psi)
                if isinstance(psi, torch.Tensor) and _scalar_outputs(ex):
                    outs = _scalar_block(outs)
                else:
                    outs = [o.to(torch.float32) if s.kind in ("zexp", "fexp") else o for s, o in zip(ex.specs, outs)]
        self._after_state(sink[0])
        if isinstance(outs, torch.Tensor):          # [1, M] block of expectation values
            out = outs[0]
        else:
            out = self._stack([o[0] for o in outs], False)
        home = self._device if self._device is not None else torch.device("cpu")
        return out.to(home) if out.device != home else out

    def execute_batch(self, params):
        """Evaluate B parameter rows ([B, P], one row = the flattened call arguments) in one launch sequence.

        Returns [B, num_measurements, ...] with an autograd graph through the adjoint kernels.
        """
        if params.dim() != 2 or params.shape[1] != self.trainable_count():
            raise ValueError(f"execute_batch expects [B, {self.trainable_count()}] parameters, "
                             f"got {tuple(params.shape)}")
        dev = _exec_device(params.device)
        rows = params.to(dev)
        if rows.is_complex():
            rows = rows.real
        if rows.dtype not in (torch.float32, torch.float64):
            rows = rows.to(torch.float32)
        ex = self.executor()
        if torch.is_grad_enabled() and rows.requires_grad:
            outs = _AdjointFunction.apply(ex, rows)
        else:
            with torch.no_grad():
                return self._run_rows(rows)[0][0]
        return outs if isinstance(outs, torch.Tensor) else self._stack(list(outs), True)

    # ------------------------------------------------------------------
    def param_shift_execute(self, *params):
        self._flatten_and_update(params)
        dev = _exec_device(self._device)
        rows = _flat_param_row(params, dev)
        out = TorchExecute.apply(self, rows)
        home = self._device if self._device is not None else torch.device("cpu")
        return out.to(home) if out.device != home else out

    def shift_recipes(self):
        """Per trainable flat index: the (coefficient, multiplier, shift) rows of its gate."""
        recipes = []
        for idx, trainable in self._trainableparams.items():
            for _ in trainable:
                recipes.append(self._operators[idx].get_parameter_shift())
        return recipes

    def jacobian_param_shift(self, rows):
        """Jacobian [num_outputs, P] at one parameter row by the shift rules, as one batched launch."""
        P = rows.shape[1]
        recipes = self.shift_recipes()
        shifted, coefs, cols = [], [], []
        for p, recipe in enumerate(recipes):
            for c, a, s in recipe:
                r = rows[0] * a
                r = r.clone()
                r[p] = r[p] + s
                shifted.append(r)
                coefs.append(c)
                cols.append(p)
        with torch.no_grad():
            batch = torch.stack(shifted).to(rows.dtype)
            out = self._run_rows(batch)[0][0].reshape(len(shifted), -1).to(torch.float64)
        jac = torch.zeros((out.shape[1], P), dtype=torch.float64, device=rows.device)
        coef = torch.tensor(coefs, dtype=torch.float64, device=rows.device)
        idx = torch.tensor(cols, dtype=torch.long, device=rows.device)
        jac.index_add_(1, idx, (out * coef[:, None]).T)
        return jac

    # ------------------------------------------------------------------
    def _after_state(self, psi):
        """``probs(qubits, after_state=True)``: post-measurement states (pytorch_backend.py:474-493), taken
        from the forward's state ``psi`` ([B, 2^n], row 0) by one gather kernel per measurement.  As in the
        reference, keys are the outcome bit strings in the order of ``qubits`` and each value is the
        normalised remaining state as a list over its first remaining axis."""
        wanted = [m for m in self._measurements if getattr(m, "after_state", False) and m.qubits is not None]
        if not wanted or not isinstance(psi, torch.Tensor):
            return
        n = self._num_qubits
        with torch.no_grad():
            for meas in wanted:
                qubits = list(meas.qubits)
                k = len(qubits)
                states = self.executor().post_measurement_states(psi[0].contiguous(), qubits)
                keys, vals = [], []
                for i in range(2 ** k):
                    sub = states[i].reshape([2] * (n - k)) if n > k else states[i].reshape(())
                    vals.append([s for s in sub] if n > k else [sub])
                    keys.append("".join(str(b) for b in dec_to_bin(i, 2 ** k)))
                self._states_after_measurement = dict(zip(keys, vals))
