"""The C-ABI library: it loads, exports every symbol include/chrono_hip.h declares, and its
host-side option and geometry functions agree with the oracle.  No GPU compute here."""
import os
import re

import numpy as np
import pytest

import chrono_photo_amd as cpa
import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "chrono_hip.h")).read()
    return set(re.findall(r"CP_API\s+[\w\s\*]+?\b(cp_\w+)\s*\(", text))


def test_header_symbols_are_exported():
    lib = cpa.load_library()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(cpa.ABI_SYMBOLS)


def test_version_and_error_string():
    lib = cpa.load_library()
    assert b"gfx950" in lib.cp_version()
    assert lib.cp_last_error() is not None


def test_threshold_matches_oracle():
    for absolute, lo, hi in [(True, 0.05, 0.2), (False, 2.0, 4.0), (True, 0.1, 0.1), (False, 0.5, 3.0)]:
        for bps in (1, 2):
            t = cpa.Threshold(absolute, lo, hi).native(bps)
            o = O.threshold(absolute, lo, hi, bps)
            assert (t.absolute, t.min, t.max) == (o.absolute, o.min, o.max)
            assert t.scale == o.scale or (np.isinf(t.scale) and np.isinf(o.scale))
            for d in np.linspace(0, 300, 61, dtype=np.float32):
                assert cpa.Threshold(absolute, lo, hi).blend_value(float(d), bps) == O.lib().or_blend_value(o, float(d))


def test_threshold_parse():
    t = cpa.Threshold.from_str("abs/0.05/0.2")
    assert t.absolute and t.min == np.float32(0.05) and t.max == np.float32(0.2)
    t = cpa.Threshold.from_str("rel/2")
    assert not t.absolute and t.min == t.max == 2.0
    with pytest.raises(ValueError):
        cpa.Threshold.from_str("foo/1")


def test_fade_matches_oracle():
    for spec in ["clamp/abs/0,0/10,1", "repeat/rel/3,1/5,0.2/9,0.7", "clamp/rel/-4,0.5/4,0.25"]:
        f = cpa.Fade.from_str(spec)
        parts = spec.split("/")
        fr = [int(p.split(",")[0]) for p in parts[2:]]
        va = [float(p.split(",")[1]) for p in parts[2:]]
        o = O.Fade(parts[0] == "repeat", parts[1] == "abs", fr, va)
        assert np.array_equal(f.values, o.table)
        for frame in range(-20, 30):
            assert f.get(frame) == o.get(frame)
    assert cpa.Fade.none().get(5) == 1.0


def test_enum_parsers():
    assert cpa.OutlierSelectionMode.from_str("forward") == cpa.OutlierSelectionMode.FORWARD
    assert cpa.BackgroundMode.from_str("median") == cpa.BackgroundMode.MEDIAN
    assert cpa.SelectionMode.from_str("darker") == cpa.SelectionMode.DARKER
    with pytest.raises(ValueError):
        cpa.OutlierSelectionMode.from_str("middle")


def test_crop_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(50):
        off = rng.integers(-8, 9, (int(rng.integers(1, 12)), 2))
        off[0] = 0
        a = cpa.Crop.create(off, 200, 150)
        b = O.crop_create(off, 200, 150)
        if b is None:
            assert a is None
        else:
            assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]


def test_video_windows_match_oracle():
    cases = [((0, 60, 1), (0, 540, 1), 600), ((0, 60, 1), (None, None, 1), 600),
             ((0, 10, 3), (0, 4, 2), 20), ((-5, 5, 2), (None, None, 3), 40),
             ((None, 7, 1), (2, None, 1), 15), ((3, None, 2), (None, 9, 1), 12)]
    for vin, vout, count in cases:
        got = cpa.video_windows(cpa.FrameRange(*vin), cpa.FrameRange(*vout), count)
        assert got == O.video_windows(vin, vout, count), (vin, vout)


def test_frame_range_parse():
    r = cpa.FrameRange.from_str("././2")
    assert (r.start, r.end, r.step) == (None, None, 2)
    r = cpa.FrameRange.from_str("0/60/.")
    assert (r.start, r.end, r.step) == (0, 60, 1)


def test_synth_host_is_deterministic():
    s = cpa.Scene(seed=3, width=40, height=30, n_frames=5, shake_max=2, texture_amp=20)
    a = s.generate()
    b = s.generate()
    assert np.array_equal(a, b)
    band = s.generate(first=1, n=2, y0=5, y1=9)
    assert np.array_equal(band, a[1:3, 5:9])
    assert s.shake_truth()[0].tolist() == [0, 0]
