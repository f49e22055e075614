"""Host-side checks of the whole-interval slicing CLI (pevsl_amd/lap_slicing.py)
that need no GPU: argument validation happens before the library loads."""
import pytest

from pevsl_amd import lap_slicing


@pytest.mark.parametrize("spec", ["0-20", "-1", "3,25", "x"])
def test_bad_slice_selection_is_rejected(spec):
    with pytest.raises(SystemExit) as e:
        lap_slicing.main(["--slices", spec, "--nslices", "20", "--quiet"])
    assert e.value.code == 2


def test_unknown_mode_is_rejected():
    with pytest.raises(SystemExit) as e:
        lap_slicing.main(["--mode", "fp16"])
    assert e.value.code == 2


@pytest.mark.parametrize("args", [["--nmat", "-5", "--gen-only"], ["--nint", "0"],
                                  ["--epochs", "0"], ["--holdout", "1.5"],
                                  ["--nmin", "10", "--nmax", "5"],
                                  ["--eval", "w.bin", "--grid-min", "50", "--grid-max", "20"]])
def test_slicer_cli_rejects_bad_sizes(args):
    """The training CLI checks its sizes before loading the library."""
    from pevsl_amd import slicer
    with pytest.raises(SystemExit) as e:
        slicer.main(args)
    assert e.value.code == 2
