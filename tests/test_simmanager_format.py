"""The native restart and energy text formats of marsmd.simmanager, without a GPU:
std::to_chars(float) shortest round-trip numbers (SimManager.cpp:14-28), the "type x y
z" restart lines and their reader."""
import numpy as np

from marsmd.simmanager import float_to_chars, ostream_double, read_restart, restart_line


def test_restart_number_format():
    # std::to_chars(float): shortest round trip, fixed unless scientific is shorter
    for v, s in [(180.5, "180.5"), (0.0, "0"), (-0.0, "-0"), (1e-5, "1e-05"), (123456789.0, "123456790"),
                 (250.0, "250"), (1e7, "1e+07"), (0.001234, "0.001234"), (-3.25, "-3.25")]:
        assert float_to_chars(v) == s, (v, float_to_chars(v), s)
    rs = np.random.default_rng(0)
    x = np.concatenate([rs.uniform(0, 2000, 20000), rs.uniform(-1, 1, 2000) * 10.0 ** rs.integers(-30, 30, 2000)])
    x = x.astype(np.float32)
    back = np.array([np.float32(float_to_chars(v)) for v in x], np.float32)
    assert np.array_equal(back.view(np.uint32), x.view(np.uint32))


def test_restart_lines_round_trip(tmp_path):
    rs = np.random.default_rng(1)
    pos = rs.uniform(0, 400, size=(100, 3)).astype(np.float32)
    types = rs.integers(0, 3, 100).astype(np.int32)
    path = tmp_path / "x.restart"
    path.write_text("".join(restart_line(t, p) for t, p in zip(types, pos)))
    t2, p2 = read_restart(path)
    assert np.array_equal(t2, types)
    assert np.array_equal(p2.view(np.uint32), pos.view(np.uint32))
    assert restart_line(2, np.array([1.5, 250.0, 1e-5], np.float32)) == "2 1.5 250 1e-05\n"


def test_energy_numbers_use_the_stream_default():
    assert ostream_double(1234567.0) == "1.23457e+06"
    assert ostream_double(0.5) == "0.5"
    assert ostream_double(-12.25) == "-12.25"
