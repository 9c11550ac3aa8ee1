"""Station output in the reference's ASCII_STATION layout.

The seismograms come off the device already in the station's wavefield
coordinate system (the rotation of Station::rotateField,
src/core/output/station-wise/station/Station.hpp:79-118, runs in the
receiver kernel; axs3d_set_receiver_frames).  This module turns them into
the reference's channels and files:

* channels (channel.hpp:39-79, solid displacement): U1, U2, U3 (the three
  components in the chosen system) and U_NORM (their Euclidean norm),
  selected by the user names "U", "U1".., "U_NORM" or "|U|", sorted by
  their standard index and de-duplicated as ChannelOptions does
  (channel.cpp:9-31);
* sampling (StationGroup::record, StationGroup.hpp:104-126): a step is
  recorded when step % sample_interval == 0 and tmin <= t <= tmax;
* files (StationIO_Ascii.cpp:17-169, StationIO.cpp:17-64), under
  <output>/stations/<group>/:
    station-centric: data_time.ascii, <key>.ascii (one row per recorded
      step, one column per channel) and list_channel.info;
    channel-centric: data_time.ascii, dir_rank0/<channel>.ascii (one row
      per recorded step, one column per station), dir_rank0/
      list_station.info and rank_station.info;
  every dump (dump_interval recorded steps, the last one shorter) is
  printed as Eigen prints a matrix: each value as std::ostream prints it
  at the stream precision (16 for time; 8 for the single-precision build,
  16 for the double-precision one), right-aligned to the widest value of
  that dump, one space between columns, then a newline.
"""
from __future__ import annotations

import os
from typing import Dict, List, Sequence

import numpy as np

# wavefield coordinate systems (channel.hpp:24); the receiver kernel's
# frame codes
WCS = {"spz": 0, "RTZ": 1, "ENZ": 2, "xyz": 3}

# standard solid displacement channels: index, name, user names, component
# (channel.hpp:41-44)
_CHANNELS = [(1, "U1", ("U1", "U"), 0), (2, "U2", ("U2", "U"), 1),
             (3, "U3", ("U3", "U"), 2), (4, "U_NORM", ("U_NORM", "|U|"), -1)]


def standard_channels(user: Sequence[str]) -> List[str]:
    """ChannelOptions (channel.cpp:9-31): every standard channel a user
    name selects, sorted by index, without repeats; an unknown name is an
    error."""
    picked = set()
    for u in user:
        found = False
        for idx, name, names, _ in _CHANNELS:
            if u in names:
                picked.add(idx)
                found = True
        if not found:
            raise ValueError("channel::solid::ChannelOptions::ChannelOptions"
                             " || Unknown channel for solid: " + u)
    return [name for idx, name, _, _ in _CHANNELS if idx in picked]


def channel_values(field: np.ndarray, channel: str) -> np.ndarray:
    """One channel of a recorded displacement [..., 3] (Station::
    computeFeedChannel, Station.hpp:120-175, displacement part)."""
    for _, name, _, comp in _CHANNELS:
        if name == channel:
            if comp >= 0:
                return field[..., comp]
            return np.sqrt((field * field).sum(axis=-1))
    raise ValueError("unknown channel " + channel)


def recorded_steps(nsteps: int, t0: float, dt: float, sample_interval=1,
                   tmin=-np.inf, tmax=np.inf) -> np.ndarray:
    """Steps StationGroup::record keeps (StationGroup.hpp:104-108)."""
    steps = np.arange(nsteps)
    t = t0 + steps * dt
    keep = (steps % sample_interval == 0) & (t >= tmin) & (t <= tmax)
    return steps[keep]


def _fmt(x: float, precision: int) -> str:
    # std::ostream << value with the default float field: printf %g
    return "%.*g" % (precision, x)


def format_block(block: np.ndarray, precision: int) -> str:
    """Eigen's matrix print (IOFormat defaults): values right-aligned to
    the widest one, ' ' between columns, '\\n' between rows, no newline
    after the last row."""
    block = np.atleast_2d(block)
    cells = [[_fmt(float(v), precision) for v in row] for row in block]
    width = max((len(c) for row in cells for c in row), default=0)
    return "\n".join(" ".join(c.rjust(width) for c in row) for row in cells)


def write_ascii_stations(output_dir: str, group: str, keys: Sequence[str],
                         user_channels: Sequence[str], times: np.ndarray,
                         data: np.ndarray, station_centric: bool = True,
                         dump_interval: int = 1000,
                         double_precision: bool = False) -> Dict[str, str]:
    """Writes one station group.  data: [sample][station][3] displacement
    in the group's wavefield coordinate system (as the device records
    it), already restricted to the recorded steps, whose times are
    `times`.  Returns the paths written, by role."""
    data = np.asarray(data, dtype=np.float64)
    times = np.asarray(times, dtype=np.float64)
    nrec, nst = data.shape[0], data.shape[1]
    if len(keys) != nst:
        raise ValueError("one key per station")
    if len(times) != nrec:
        raise ValueError("one time per recorded sample")
    channels = standard_channels(user_channels)
    prec_wave = 16 if double_precision else 8
    if not double_precision:
        # the single-precision build buffers Real = float
        data = data.astype(np.float32).astype(np.float64)
    gdir = os.path.join(output_dir, "stations", group)
    os.makedirs(gdir, exist_ok=True)
    paths = {}
    vals = np.stack([channel_values(data, c) for c in channels], axis=2)
    if not double_precision:
        vals = vals.astype(np.float32).astype(np.float64)
    dump = max(1, min(dump_interval, nrec))
    time_path = os.path.join(gdir, "data_time.ascii")
    paths["data_time"] = time_path
    if station_centric:
        with open(os.path.join(gdir, "list_channel.info"), "w") as f:
            for c in channels:
                f.write(c + "\n")
        paths["list_channel"] = os.path.join(gdir, "list_channel.info")
        outs = {k: open(os.path.join(gdir, k + ".ascii"), "w") for k in keys}
        for k in keys:
            paths["station:" + k] = os.path.join(gdir, k + ".ascii")
    else:
        rdir = os.path.join(gdir, "dir_rank0")
        os.makedirs(rdir, exist_ok=True)
        with open(os.path.join(rdir, "list_station.info"), "w") as f:
            for k in keys:
                f.write(k + "\n")
        with open(os.path.join(gdir, "rank_station.info"), "w") as f:
            f.write("MPI_RANK STATION_KEY STATION_INDEX_IN_RANK\n")
            for i, k in enumerate(keys):
                f.write(f"0 {k} {i}\n")
        paths["list_station"] = os.path.join(rdir, "list_station.info")
        paths["rank_station"] = os.path.join(gdir, "rank_station.info")
        outs = {c: open(os.path.join(rdir, c + ".ascii"), "w")
                for c in channels}
        for c in channels:
            paths["channel:" + c] = os.path.join(rdir, c + ".ascii")
    try:
        with open(time_path, "w") as ft:
            for b0 in range(0, nrec, dump):
                b1 = min(nrec, b0 + dump)
                ft.write(format_block(times[b0:b1, None], 16) + "\n")
                if station_centric:
                    for i, k in enumerate(keys):
                        outs[k].write(format_block(vals[b0:b1, i, :],
                                                   prec_wave) + "\n")
                else:
                    for j, c in enumerate(channels):
                        outs[c].write(format_block(vals[b0:b1, :, j],
                                                   prec_wave) + "\n")
    finally:
        for f in outs.values():
            f.close()
    return paths


def read_ascii_block(path: str) -> np.ndarray:
    """The numbers of an ASCII station file, one row per recorded step."""
    rows = [ln.split() for ln in open(path) if ln.strip()]
    return np.array([[float(v) for v in r] for r in rows])
