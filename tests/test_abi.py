"""CPU-side checks of the C ABI boundary: the library loads and exports
every entry point include/bruno_hip.h declares (no compute call)."""
import os
import re

import bruno_amd
from helpers import ROOT


def declared():
    with open(os.path.join(ROOT, "include", "bruno_hip.h")) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(bruno_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = bruno_amd.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(bruno_amd.EXPORTS) == names


def test_status_strings_and_bad_input():
    L = bruno_amd.lib()
    assert L.bruno_status_string(0) == b"ok"
    assert L.bruno_status_string(3).startswith(b"capacity")
    # NULL descriptor is rejected before any device call
    assert L.bruno_job_create(None, None) == 1


def test_filter_modes_on_the_host():
    """thresholding 0 / 1 / 2 (plain, k2 filter, k+1-mer edge filter); other
    values are rejected before any device call."""
    import ctypes
    import pytest
    from bruno_amd.pipeline import job_desc, thresholds_L
    assert job_desc(31).thresholding == 0
    assert job_desc(31, thresholds_L(2)).thresholding == 1
    d = job_desc(31, thresholds_L(2), edge_filter=True)
    assert d.thresholding == 2 and d.thresholds[2] == 2 and d.thresholds[3] == 65536
    with pytest.raises(ValueError):
        job_desc(31, None, edge_filter=True)
    h = ctypes.c_void_p()
    for bad in (-1, 3):
        d.thresholding = bad
        assert bruno_amd.lib().bruno_job_create(ctypes.byref(d), ctypes.byref(h)) == 1


def test_host_generator_is_deterministic():
    from bruno_amd.configs import config2
    c = config2(seed=4, scale=0.0005)["synth"]
    a = bruno_amd.synth_reads_host(c)
    b = bruno_amd.synth_reads_host(c)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    assert set(bytes(a[0])) <= set(b"ACGT")


def test_tuning_knobs_validated_on_the_host():
    from bruno_amd.autotune import get_tuning, set_tuning
    import pytest
    before = get_tuning()
    for bad in ({"parse_block": 192}, {"ruler_sample": 100}, {"ruler_sample": 4}, {"parse_grid": 1},
                {"count_regions": 3}, {"region_slots": 2048}, {"count_groups": 65}):
        with pytest.raises(bruno_amd.BrunoError):
            set_tuning(**bad)
    assert get_tuning() == before
    assert before["count_groups"] == 0  # automatic
    set_tuning(parse_block=512, ruler_sample=128, count_regions=2, region_slots=1280, count_groups=64)
    t = get_tuning()
    assert (t["parse_block"], t["ruler_sample"], t["count_regions"], t["region_slots"], t["count_groups"]) == \
        (512, 128, 2, 1280, 64)
    set_tuning(**before)
