"""Frame files of a run and the readers a restart needs.

The library writes config_<t>_.txt / quats_<t>_.txt / force_<t>_.txt / brownian_stresslet_<t>_.txt
(src/write_configuration.cpp:10-166; ps_write_outputs).  The readers follow the reference's analysis
scripts (test_sim_dumbbells_thermal.py:189-214): frames sorted by the time in the file name, one row
per particle.  A restart (test_sim_dumbbells_thermal.py:541-584) builds PolyStokes(..., restart=t_last)
and passes the last position and quaternion frames to initial_configuration.
"""
import ctypes as C
import os

import numpy as np

from .polystokes import lib

PREFIXES = ("config_", "quats_", "force_", "brownian_stresslet_")


def format_time(t):
    """The <t> text of a frame file name: a default std::ostream's output (6 significant digits)."""
    buf = C.create_string_buffer(64)
    rc = lib().ps_format_time(float(t), buf, 64)
    if rc != 0:
        raise ValueError(f"cannot format {t!r}")
    return buf.value.decode()


def frame_path(output_dir, prefix, t):
    return os.path.join(output_dir, f"{prefix}{format_time(t)}_.txt")


def read_frames(output_dir, prefix="config_"):
    """(times, frames) of every <prefix><t>_.txt in output_dir, sorted by time; frames has shape
    (nframes, rows, cols)."""
    files = [f for f in os.listdir(output_dir) if f.startswith(prefix) and f.endswith("_.txt")]
    # the time is the text between the prefix and "_.txt"
    times = [float(f[len(prefix):-len("_.txt")]) for f in files]
    order = np.argsort(times, kind="stable")
    frames = [np.loadtxt(os.path.join(output_dir, files[i]), ndmin=2) for i in order]
    return np.asarray(times)[order], np.asarray(frames)


def read_trajectory(output_dir):
    """Position frames (nframes, Np, 3) and their times."""
    return read_frames(output_dir, "config_")


def read_quaternions(output_dir):
    """Quaternion frames (nframes, Nc, 4) and their times."""
    return read_frames(output_dir, "quats_")
