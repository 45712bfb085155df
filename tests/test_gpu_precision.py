"""The precision study (continuous Au 30 keV, three Bethe laws, two tilts; DS
Cu 20 keV): FP32 against the all-FP64 GPU run of the same 2^22 electrons per
workload, every output against its bar, and the table printed."""
import pytest

pytestmark = pytest.mark.gpu


def test_precision_study_every_output_within_its_bar():
    from mcd import precision
    rows = precision.run_study(1 << 22)
    print("\n" + precision.markdown(rows))
    for name, c in rows:
        assert c["identical"] >= 0.999, (name, c)
        assert c["relE_max"] <= 1e-5, (name, c)
        assert c["dir_max"] <= 1e-5, (name, c)
        assert c["pos_max"] <= 1e-5, (name, c)
        assert abs(c["y32"] - c["y64"]) <= 3 * c["y_sigma"], (name, c)
