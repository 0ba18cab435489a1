"""The GPU path against the reference's own recorded output: the recorded
Taylor-Green runs at 64^3 and 128^3 (taylorgreen_vortex/taylorgreen_run64,
taylorgreen_run128; settings in tests/test_oracle_golden.py), two RK4 steps
in FP64 through the C ABI, and cs3d_diag_taylor_green against the first two
rows of their taylorgreen.out (tests/golden/) at the oracle's bounds."""
import pytest

import cs3d
from test_oracle_golden import check_row, recorded_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [64, 128])
def test_gpu_tgv_recorded_rows(N):
    kw, ic, rows = recorded_case(N)
    g = cs3d.Solver(cs3d.config(N, N, N, precision=cs3d.FP64, **kw))
    g.set_initial_conditions(*ic)
    g.set_diagnostics(True)
    for r in range(2):
        g.step(1)
        d = g.diag_taylor_green()
        print(f"{N}^3 step {r + 1}: gpu {d}\n            recorded {rows[r]}", flush=True)
        check_row(d, rows[r])
    g.close()
