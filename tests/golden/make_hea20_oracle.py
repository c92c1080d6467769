"""Write the complex128 oracle's values and gradients for two parameter rows of the
20-qubit, 10-layer hardware-efficient benchmark batch (tests/golden/hea20_d10_rows.json).

The batch is the benchmark's: ``numpy.random.default_rng(1).uniform(0, 2 pi, (256, 400))``,
parameters ordered layer -> qubit -> (RY theta, RZ phi).  For rows 0 and 31 the file holds
the parameters themselves, the 20 values <Z_i> and the 400 derivatives of sum_i <Z_i>,
all from the oracle (oracle/sv_oracle.py) in complex128 with torch autograd.  One row takes
about a minute on 8 host cores, which is why the GPU parity test reads this file instead
of running the oracle.

Run from the repository root:  python tests/golden/make_hea20_oracle.py
"""

import json
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import sv_oracle as O  # noqa: E402

N, LAYERS, BATCH, SEED = 20, 10, 256, 1
ROWS = (0, 31)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hea20_d10_rows.json")


def benchmark_batch():
    return np.random.default_rng(SEED).uniform(0, 2 * np.pi, size=(BATCH, LAYERS * N * 2))


def oracle_row(theta_row):
    t = torch.tensor(theta_row.reshape(LAYERS, N, 2), dtype=torch.float64, requires_grad=True)
    ref = O.execute(N, O.hea_gates(N, t)[0], [("expval", ("PauliZ", [q])) for q in range(N)])
    ref.sum().backward()
    return ref.detach().numpy(), t.grad.reshape(-1).numpy()


def main():
    theta = benchmark_batch()
    rows = []
    for i in ROWS:
        vals, grads = oracle_row(theta[i])
        rows.append({"row": i, "params": theta[i].tolist(), "values": vals.tolist(), "grad_of_sum": grads.tolist()})
        print(f"row {i}: done", flush=True)
    doc = {"_about": "complex128 oracle (oracle/sv_oracle.py) on the 20-qubit 10-layer hardware-efficient "
                     "ansatz: <Z_i> for i = 0..19 and d(sum_i <Z_i>)/d(params); batch "
                     "numpy.random.default_rng(1).uniform(0, 2 pi, (256, 400)); written by "
                     "tests/golden/make_hea20_oracle.py",
           "qubits": N, "layers": LAYERS, "rows": rows}
    with open(OUT, "w") as f:
        json.dump(doc, f)


if __name__ == "__main__":
    main()
