"""Command line: see ppmpa_amd.cli."""
import sys

from .cli import pm_main

if __name__ == "__main__":
    sys.exit(pm_main())
