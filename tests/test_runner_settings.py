"""RunSettings (src/runner.rs:29-82) as mirrored by the library's host code; the reference's own
runner tests (src/runner.rs:287-316)."""
import pytest


def test_defaults_match_the_current_run(treepm):
    d = treepm.RunSettings.default()
    assert d.particles_per_side == 40
    assert d.z_start == 500.0
    assert d.z_end == pytest.approx(-0.9)
    assert d.treepm == 1
    assert d.steps == 512 and d.seed == 1
    assert d.theta == 0.5 and d.softening_scale == pytest.approx(0.05)


def test_sanitize_keeps_defaults_and_clamps_bad_values(treepm):
    d = treepm.RunSettings.default()
    assert d.sanitized().as_dict() == d.as_dict()
    bad = treepm.RunSettings.default()
    bad.particles_per_side = 500
    bad.steps = 1
    bad.z_start = 0.0
    bad.z_end = 5.0
    bad.theta = 9.0
    s = bad.sanitized()
    assert s.particles_per_side == 200
    assert s.steps == 16
    assert s.z_start == 10.0
    assert s.z_end < s.z_start
    assert s.theta == 1.0


def test_sanitize_lower_clamps_and_rejects_non_finite(treepm):
    s = treepm.RunSettings.default()
    s.particles_per_side = 1
    s.z_start = 5000.0
    s.z_end = -2.0
    s.theta = 0.0
    s.softening_scale = 3.0
    s.pacing_ms = 10 ** 6
    t = s.sanitized()
    assert (t.particles_per_side, t.z_start, t.theta, t.pacing_ms) == (5, 2000.0, pytest.approx(0.1), 2000)
    assert t.z_end == pytest.approx(-0.95) and t.softening_scale == 0.5
    s.z_start = float("nan")
    with pytest.raises(treepm.TpmError):
        s.sanitized()
