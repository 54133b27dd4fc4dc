"""``python -m pyising_amd`` runs the command line interface (pyising_amd.cli)."""

import sys

from .cli import main

sys.exit(main())
