"""`python marl-ochre_amd [flags]`: the command-line run (see run.py), the
counterpart of the reference's `marl-engine` binary."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from __graft_entry__ import load_package  # noqa: E402

if __name__ == "__main__":
    load_package()
    from marl_ochre_amd import run  # noqa: E402

    sys.exit(run.main())
