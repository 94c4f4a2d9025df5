"""``python -m dxtb_amd``: the command line single point (dxtb_amd.cli)."""

import sys

from .cli import main

sys.exit(main())
