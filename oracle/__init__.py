"""CPU oracle of the path tracer — test infrastructure only.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as
the checker; the product package (trayracer_amd) never imports it.
"""
from .oracle_py import Oracle, build_oracle, load_oracle  # noqa: F401
