"""Command line: see ppmpa_amd.cli."""
import sys

from .cli import rtc_main

if __name__ == "__main__":
    sys.exit(rtc_main())
