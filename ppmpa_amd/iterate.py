"""Command line: see ppmpa_amd.cli."""
import sys

from .cli import iterate_main

if __name__ == "__main__":
    sys.exit(iterate_main())
