"""Pass manager factory (qgpusim/transpiler/pm.py:11-19).

``make_custom_pm(nL)`` returns ``(pm, lqr_pass)``; ``pm.run(circuit)`` returns
the circuit re-emitted in tile-schedule order and leaves the plan in
``lqr_pass.property_set["flat_tiling_plan"]`` (qgpusim/runner.py:48-50).
"""
from ..circuit import circuit_to_dag, dag_to_circuit
from .passes import LazyQubitReordering


class PassManager:
    def __init__(self, passes):
        self.passes = list(passes)

    def run(self, circuit):
        dag = circuit_to_dag(circuit)
        for p in self.passes:
            dag = p.run(dag)
        return dag_to_circuit(dag)


def make_custom_pm(nL: int):
    lqr = LazyQubitReordering(nL)
    return PassManager([lqr]), lqr
