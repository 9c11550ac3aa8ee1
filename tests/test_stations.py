"""Station files in the reference's ASCII_STATION layout
(StationIO_Ascii.cpp:17-169, StationIO.cpp:17-64) and the channel rules
of channel.cpp:9-31 / channel.hpp:39-79, on the host.

The value formatting is pinned against what C++ streams and Eigen's
matrix printer produce by definition: std::ostream << double with
precision p and the default float field is printf's %.{p}g; Eigen's
IOFormat defaults right-align every value to the widest one of the
matrix, one space between columns, '\\n' between rows."""
import numpy as np
import pytest

from axisem3d_amd import stations


def test_ostream_precision_is_printf_g():
    assert stations._fmt(1.0, 8) == "1"
    assert stations._fmt(-2.5, 8) == "-2.5"
    assert stations._fmt(1e-5, 8) == "1e-05"
    assert stations._fmt(123456789.0, 8) == "1.2345679e+08"
    assert stations._fmt(0.1, 16) == "0.1"
    assert stations._fmt(1.0 / 3.0, 16) == "0.3333333333333333"
    assert stations._fmt(float(np.float32(0.1)), 8) == "0.1"
    assert stations._fmt(float(np.float32(0.1)), 16) == "0.1000000014901161"


def test_matrix_block_is_right_aligned_to_the_widest_value():
    b = np.array([[1.0, -2.5], [1e-5, 123456789.0]])
    txt = stations.format_block(b, 8)
    assert txt == ("            1          -2.5\n"
                   "        1e-05 1.2345679e+08")
    # a column (the time file) and no trailing newline
    assert stations.format_block(np.array([[0.5], [10.25]]), 16) == \
        "  0.5\n10.25"


def test_channel_selection_follows_the_standard_order():
    assert stations.standard_channels(["U"]) == ["U1", "U2", "U3"]
    assert stations.standard_channels(["|U|", "U3", "U1", "U"]) == \
        ["U1", "U2", "U3", "U_NORM"]
    assert stations.standard_channels(["U_NORM", "U2"]) == ["U2", "U_NORM"]
    with pytest.raises(ValueError, match="Unknown channel for solid: UR"):
        stations.standard_channels(["UR"])


def test_channel_values_and_norm():
    f = np.array([[3.0, 4.0, 12.0], [1.0, 0.0, 0.0]])
    assert np.array_equal(stations.channel_values(f, "U2"), [4.0, 0.0])
    assert np.allclose(stations.channel_values(f, "U_NORM"), [13.0, 1.0])


def test_recorded_steps_interval_and_window():
    # StationGroup::record: step % interval == 0 and tmin <= t <= tmax
    st = stations.recorded_steps(20, t0=-1.0, dt=0.25, sample_interval=3,
                                 tmin=0.0, tmax=3.0)
    # t = -1 + 0.25 step lies in [0, 3] for steps 4..16
    assert list(st) == [6, 9, 12, 15]


def _data(nrec=7, nst=3, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nrec, nst, 3)) * 10.0 ** rng.integers(
        -12, -6, (nrec, nst, 3))


@pytest.mark.parametrize("double", [False, True])
def test_station_centric_files(tmp_path, double):
    data = _data()
    t = -1.0 + 0.1 * np.arange(7)
    keys = ["II.AAK", "II.ABC", "IU.XYZ"]
    p = stations.write_ascii_stations(str(tmp_path), "grp", keys,
                                      ["U", "U_NORM"], t, data,
                                      station_centric=True, dump_interval=3,
                                      double_precision=double)
    gdir = tmp_path / "stations" / "grp"
    assert (gdir / "list_channel.info").read_text() == \
        "U1\nU2\nU3\nU_NORM\n"
    # three dumps (3 + 3 + 1 rows), each block ends with a newline
    txt = (gdir / "II.ABC.ascii").read_text()
    assert txt.count("\n") == 7
    vals = stations.read_ascii_block(p["station:II.ABC"])
    assert vals.shape == (7, 4)
    ref = data[:, 1, :]
    if not double:
        ref = ref.astype(np.float32).astype(np.float64)
    rtol = 1e-15 if double else 1e-7
    assert np.allclose(vals[:, :3], ref, rtol=rtol, atol=0)
    assert np.allclose(vals[:, 3], np.linalg.norm(ref, axis=1), rtol=2e-7)
    tt = stations.read_ascii_block(p["data_time"])[:, 0]
    assert np.array_equal(tt, t)


def test_channel_centric_files(tmp_path):
    data = _data(nrec=4, nst=2)
    t = np.arange(4) * 0.5
    keys = ["AA.S1", "AA.S2"]
    p = stations.write_ascii_stations(str(tmp_path), "g2", keys, ["U3"], t,
                                      data, station_centric=False)
    gdir = tmp_path / "stations" / "g2"
    assert (gdir / "rank_station.info").read_text() == \
        ("MPI_RANK STATION_KEY STATION_INDEX_IN_RANK\n"
         "0 AA.S1 0\n0 AA.S2 1\n")
    assert (gdir / "dir_rank0" / "list_station.info").read_text() == \
        "AA.S1\nAA.S2\n"
    vals = stations.read_ascii_block(p["channel:U3"])
    assert vals.shape == (4, 2)
    ref = data[:, :, 2].astype(np.float32).astype(np.float64)
    assert np.allclose(vals, ref, rtol=1e-7, atol=0)


# ---- station frames: the oracle's restatement of Station::rotateField,
# pinned by geometry (a station at colatitude theta from the source on a
# spherical mesh; back azimuth baz = direction from the station to the
# source, clockwise from north)
def _one(v):
    return np.asarray(v, dtype=np.float64).reshape(1, 1, 3)


def test_rtz_maps_the_local_vertical_and_the_great_circle_direction():
    import oracle
    th = 0.7
    up = _one([np.sin(th), 0.0, np.cos(th)])            # radial outward
    away = _one([np.cos(th), 0.0, -np.sin(th)])         # increasing theta
    phi_hat = _one([0.0, 1.0, 0.0])
    rot = lambda v: oracle.rotate_station(v, [1], [th], [0.0], [0.0], False)
    assert np.allclose(rot(up), _one([0, 0, 1]), atol=1e-15)
    assert np.allclose(rot(away), _one([1, 0, 0]), atol=1e-15)
    assert np.allclose(rot(phi_hat), _one([0, 1, 0]), atol=1e-15)


def test_enz_points_north_towards_the_source():
    import oracle
    th = 0.4
    away = _one([np.cos(th), 0.0, -np.sin(th)])
    # back azimuth 0: the source lies due north, so "away" is south
    enz = oracle.rotate_station(away, [2], [th], [0.0], [0.0], False)
    assert np.allclose(enz, _one([0, -1, 0]), atol=1e-15)
    # back azimuth 90 deg: the source lies due east, "away" is west
    enz = oracle.rotate_station(away, [2], [th], [np.pi / 2], [0.0], False)
    assert np.allclose(enz, _one([-1, 0, 0]), atol=1e-15)


def test_xyz_undoes_the_azimuth_and_spz_is_the_identity():
    import oracle
    phi = 1.1
    # the Cartesian x axis seen from azimuth phi: (cos phi, -sin phi, 0)
    x_spz = _one([np.cos(phi), -np.sin(phi), 0.0])
    for cart in (True, False):
        xyz = oracle.rotate_station(x_spz, [3], [0.3], [0.2], [phi], cart)
        assert np.allclose(xyz, _one([1, 0, 0]), atol=1e-15)
    v = np.random.default_rng(0).standard_normal((4, 2, 3))
    assert np.array_equal(oracle.rotate_station(v, [0, 0], [1, 2], [3, 4],
                                                [5, 6], False), v)
    # Cartesian meshes: RTZ is SPZ
    assert np.array_equal(oracle.rotate_station(v, [1, 1], [1, 2], [3, 4],
                                                [5, 6], True), v)
    # every frame keeps lengths
    w = oracle.rotate_station(v, [2, 3], [0.3, 1.2], [0.1, 2.0],
                              [0.5, 4.0], False)
    assert np.allclose(np.linalg.norm(w, axis=2), np.linalg.norm(v, axis=2))
