"""Command line: see ppmpa_amd.cli."""
import sys

from .cli import ppmpa_main

if __name__ == "__main__":
    sys.exit(ppmpa_main())
