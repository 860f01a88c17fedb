"""Oracle package: a float64 NumPy restatement of jax-power's mesh power-spectrum path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` (as the checker)
and ``bench.py``'s ``cpu_baseline`` leg may import it; the shipped package never does.
"""
from . import power_oracle  # noqa: F401
