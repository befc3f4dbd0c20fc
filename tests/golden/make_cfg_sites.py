"""Writes tests/golden/cfg/<program>.sites: the HipSite lines of the example
programs (tests/golden/cfg_programs.py), printed as GHC's derived Show prints
them.  The build compiles these files with hydgen; a CPU test checks that the
committed files equal this script's output.

    python tests/golden/make_cfg_sites.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import cfg_programs  # noqa: E402
from cfgshow import show  # noqa: E402

OUT = os.path.join(HERE, "cfg")


def render(name):
    recs = cfg_programs.sites(name) + cfg_programs.groups(name)
    return "".join(show(s) + "\n" for s in recs)


def main():
    os.makedirs(OUT, exist_ok=True)
    for name in cfg_programs.PROGRAMS:
        with open(os.path.join(OUT, name + ".sites"), "w") as f:
            f.write(render(name))
    print("wrote", len(cfg_programs.PROGRAMS), "site files to", OUT)


if __name__ == "__main__":
    main()
