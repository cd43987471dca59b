"""`python -m lensed_amd`: the command line of lensed_amd.cli."""
import sys

from .cli import main

sys.exit(main())
