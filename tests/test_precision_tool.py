"""The per-bin error measure of the FP32-vs-FP64 precision run
(gabe_amd/precision.py): the worst relative error over the bins above 1e-9 of
each column's maximum, bin 0 excluded, for the field and GW columns."""
import numpy as np

from gabe_amd import precision


def test_bin_errors_takes_the_live_bins_only():
    ref = np.ones((6, 8))
    ref[:, 5] = 1e-12          # dead bin: below 1e-9 of the maximum
    got = ref.copy()
    got[0, 0] = 5.0            # bin 0 is excluded
    got[0, 5] = 1.0            # error in a dead bin is ignored
    got[2, 3] = 1.0 + 2e-4     # second field column
    got[4, 6] = 1.0 - 3e-5     # GW column
    e = precision.bin_errors(got, ref)
    assert abs(e["field"] - 2e-4) < 1e-12
    assert abs(e["gw"] - 3e-5) < 1e-12


def test_bin_errors_without_gw_column():
    ref = np.ones((4, 5))
    got = ref * (1 + 1e-3)
    e = precision.bin_errors(got, ref)
    assert abs(e["field"] - 1e-3) < 1e-12 and e["gw"] == 0.0
