"""CPU oracle for the LBGK hot path — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package.  It wraps ``liblbgk_oracle.so`` (built from
``lbgk_oracle.cpp`` by ``oracle/Makefile``), a C++ restatement of ASL's
``asl::LBGK`` pool/window streaming, collision and LBGK boundary conditions
(see the header of ``lbgk_oracle.cpp`` for the ASL files each part follows).
"""
from .lbgk import OracleLattice, df_map, load_library, surface_distance_map, template, triangle_d2  # noqa: F401
