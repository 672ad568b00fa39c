#!/usr/bin/env python3
"""Drop-in for the reference's fibers_runner.py (run / validate / benchmark)
on the MI355X HIP path; see rigidfibers_amd/runner.py."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from rigidfibers_amd.runner import main  # noqa: E402

if __name__ == "__main__":
    sys.exit(main())
