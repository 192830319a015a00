"""The HBM read-bandwidth probe (stc_probe_read_bandwidth): a measurement hook,
not a count, so it is checked for sane, ordered figures only."""
import pytest

import signed_triangle_counting_amd as stc

pytestmark = pytest.mark.gpu


def test_probe_read_bandwidth_is_positive_and_below_peak():
    r = stc.probe_read_bandwidth(0)
    print("probe", r, flush=True)
    assert set(r) == {"sequential", "segment_128B", "segment_256B", "segment_512B",
                      "segment_1024B"}
    for k, v in r.items():
        assert 0.1 < v < 8.5, (k, v)  # MI355X data-sheet peak: 8 TB/s
    # longer segments use more of each 128-byte line fetched: never much slower
    assert r["segment_1024B"] > 0.8 * r["segment_128B"]
