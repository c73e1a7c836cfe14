#!/usr/bin/env python
"""`pgm-jax md` on the MI355X: rigid pGM-water MD from Amber inputs (see pgm-jax_amd/driver.py).

    python pgm-jax_amd/run_md.py -p examples/water512/pgm3p25_512.prmtop -c examples/water512/pgm3p25_512.rst7 \
        -o runs/md --thermostat none --barostat none --nsteps 10 --report-every 5 --precision double
"""

import importlib.util
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _package():
    if "pgm_jax_amd" in sys.modules:
        return sys.modules["pgm_jax_amd"]
    spec = importlib.util.spec_from_file_location("pgm_jax_amd", os.path.join(HERE, "__init__.py"), submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pgm_jax_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


if __name__ == "__main__":
    sys.exit(_package().driver.main(sys.argv[1:]))
