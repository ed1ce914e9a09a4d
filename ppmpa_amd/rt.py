"""Command line: see ppmpa_amd.cli."""
import sys

from .cli import rt_main

if __name__ == "__main__":
    sys.exit(rt_main())
