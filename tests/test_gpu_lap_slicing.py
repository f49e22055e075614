"""GPU test of the whole-interval slicing CLI (pevsl_amd/lap_slicing.py) on
the 20^3 Laplacian over [0.6, 1.2] in 5 slices: every mode finds exactly the
closed-form count in every slice, eigenvalues to 1e-10 relative, and the
summary's relative residual is the absolute one divided by |lambda|.

The mixed modes store the basis in FP32, so their residuals are bounded by
1e-6 * max(|lmin|, |lmax|) = 1.2e-5 (lmax = 12) rather than by 1e-8.
"""

import json

import pytest

from pevsl_amd import lap_slicing

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,wave", [("fp64", 0), ("mixed", 0), ("mixed", 2),
                                       ("mixed-mlp", 0)])
def test_lap_slicing_cli_counts_exact(mode, wave, capsys):
    print("\n[gpu case] lap_slicing 20^3 [0.6,1.2] 5 slices mode %s wave %d" % (mode, wave),
          flush=True)
    args = ["--grid", "20", "--interval", "0.6", "1.2", "--nslices", "5", "--mode", mode,
            "--wave", str(wave), "--quiet"]
    rc = lap_slicing.main(args)
    line = capsys.readouterr().out.strip().splitlines()[-1]
    out = json.loads(line)
    assert rc == 0
    assert out["slices_not_exact"] == []
    assert out["found"] == out["exact"] > 0
    assert len(out["per_slice"]) == 5
    assert out["max_rel_err"] < 1e-10
    res_tol = 1e-8 if mode == "fp64" else 1.2e-5
    assert out["max_residual"] < res_tol
    # accepted Ritz values lie in [0.6, 1.2] up to the acceptance margin
    assert out["max_residual"] / 1.2 * (1 - 1e-3) <= out["max_rel_residual"]
    assert out["max_rel_residual"] <= out["max_residual"] / 0.6 * (1 + 1e-3)
    if mode == "mixed-mlp":
        assert out["judge"] is not None
