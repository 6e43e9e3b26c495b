"""The ``fdtd_3d_bin`` command (the reference programs' argv / file contract).

``fdtd_3d_bin N_FRAMES STEPS_PER_FRAME SNAPSHOT_COUNT [E_KINETIC | -lambda] [PARAMS_FILE]``
reads ./potential.bin and writes ./snapshots.bin like ``fdtd_3d.cu:275-403``
(a PARAMS_FILE with an NZ line) or ``fdtd.cu:323-552`` (2D, no NZ line); its
source is ``csrc/fdtd_3d_main.cpp`` and it calls ``libwapfdtd.so``.
"""
import os
import shutil

from .build import CLI_PATH, build_cli


def ensure_cli_binary(dst=None):
    """Make sure the command exists (building it if needed) and copy it to ``dst``."""
    if not os.path.exists(CLI_PATH):
        build_cli()
    if dst is not None and os.path.abspath(dst) != CLI_PATH:
        shutil.copyfile(CLI_PATH, dst)
        os.chmod(dst, 0o755)
    return dst or CLI_PATH
