#!/usr/bin/env python3
"""Experiment driver with the reference's CLI contract (main.py run|verify|bench,
--config, --set), running the v0.5 ocean-core step on the GPU; see
hydrostatic_ocean_core_benchmark_amd/cli.py."""

import sys

from hydrostatic_ocean_core_benchmark_amd.cli import main

if __name__ == "__main__":
    sys.exit(main())
