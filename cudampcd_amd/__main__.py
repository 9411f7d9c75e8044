"""``python -m cudampcd_amd <input_file> [--seed S]``: run a parameter file."""
import sys

from .cli import main

sys.exit(main())
