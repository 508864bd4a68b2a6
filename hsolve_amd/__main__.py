"""python -m hsolve_amd ... runs the command line (hsolve_amd.cli)."""
import sys
import warnings

from .cli import main
from .errors import HsolveWarning

warnings.simplefilter("ignore", HsolveWarning)  # the exit status carries success
sys.exit(main())
