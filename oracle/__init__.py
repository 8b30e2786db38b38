"""CPU oracle for the statevector tile path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product package imports this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker (or as the timed CPU port), never as the thing measured
or shipped.

Contents and what pins each part (a pin is an independent check that a
plausible mistake would fail):

* ``gates``    -- closed-form gate matrices in Qiskit's little-endian convention.
                  Pinned by textbook identities (H^2=I, S^2=Z, T^2=S,
                  RX(pi)=-iX, ...) and by the little-endian CX action.
* ``dense``    -- FP64 tensor-contraction statevector simulator.  Pinned against
                  a brute-force 2^n x 2^n operator built index by index for n<=8.
* ``permute``  -- the layout -> bit-map restatement of
                  qgpusim_cuda/src/bindings.cpp:447-464 and the scatter of
                  qgpusim_cuda/src/kernels.cu:176-184.  Pinned against
                  numpy.transpose of the (2,)*n tensor.
* ``lqr``      -- literal restatement of the flat time-space tiling
                  (qgpusim/transpiler/passes.py:29-37, 89-357, 359-398).  Pinned
                  by the reference-held schedule counts in results/result.csv
                  (rows 6 and 13: 5 tiles/4 reorders/624 ops at n=20 d=16 nL=16;
                  9/8/784 at n=25), by schedule invariants, and by brute-force
                  equality of the tiled run with program-order simulation.
* ``circuits`` -- seeded generators for the five benchmark workloads.  QFT is
                  pinned by numpy.fft (the circuit equals the unitary DFT).
* ``sv_ref.c`` -- plain-C per-gate simulator (FP64 checker, FP32 CPU port)
                  restating the pair/quad/permute index math of
                  qgpusim_cuda/src/kernels.cu:30-49, 83-136, 173-184.  Pinned
                  against ``dense`` and ``permute``.

The reference's published rows (results/result.csv, kept as
tests/golden/result_csv_custom.json) also fix the shape of ``transpile``'s
output on its random circuits: one single-qubit gate per qubit and layer.
The rest of the Qiskit-only pieces (``DAGCircuit.topological_op_nodes``
order, the rest of ``transpile``) is unpinned: Qiskit is not installed, and
the oracle defines the gate order as program order.
"""
