"""The learned closure at reduced size: a coarse 32^3 D3Q19 lattice against the
128^3 run it coarsens (decaying turbulence, fine tau 0.505, coarse tau
0.50125).  Plain LBGK is at the edge of stability there; the closure, trained
on one realisation of the initial field and evaluated on another, must stay
finite, follow the fine run's kinetic-energy decay better than plain LBGK
does, and run much faster than the fine lattice.

The same study at 128^3 against 512^3 is ``bench.py --closure-study 128
--tau-fine 0.505``.
"""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_closure_beats_plain_at_32():
    from asl_amd import closure
    r = closure.study(32, 0.505, chunk=50, snapshots=12, eval_chunks=16, train_steps=1500, batch=8192,
                      log=lambda m: print(m, flush=True))
    print({k: r[k] for k in ("output_scale_factor", "max_rel_ke_decay_error_plain",
                             "max_rel_ke_decay_error_closure", "speedup_closure_vs_fine")})
    assert r["train_loss_last"] < r["train_loss_first"]
    assert r["output_scale_factor"] > 0.0  # calibration keeps the closure on
    ec, ep = r["max_rel_ke_decay_error_closure"], r["max_rel_ke_decay_error_plain"]
    assert math.isfinite(ec) and all(math.isfinite(v) for v in r["ke_coarse_closure"])
    assert ec < 0.3
    assert ep > 1.5 * ec  # plain LBGK: blown up (inf) or clearly worse
    assert r["speedup_closure_vs_fine"] > 4.0
