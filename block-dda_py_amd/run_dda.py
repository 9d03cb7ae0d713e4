"""Sweep driver: drop-in for the reference's ``run_dda.py``.

Fills every not-yet-computed condition of an existing results file (created
like ``dda_results/create_h5py.ipynb`` does, or with
``bdda_amd.sweep.create_store``), with the reference's settings: GRE shapes
from RNG_SEED, orientations from RNG_SEED + 1, up to MAX_TRY tries with
fresh orientations, NaN fill, Mie of the volume-equivalent sphere.  The
solves run on the GPU (see ``bdda_amd.sweep``).  An .npz results file is
used when h5py is not installed.

    python run_dda.py [--output dda_results/pcas_ocbs_simulated_data.hdf5]
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from bdda_amd import sweep  # noqa: E402

RNG_SEED = sweep.RNG_SEED   # GRE shape uses RNG_SEED; Euler angles use RNG_SEED + 1
MAX_TRY = sweep.MAX_TRY     # max DDA retries when the solver does not converge
OUTPUT_FILE = "dda_results/pcas_ocbs_simulated_data.hdf5"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--output", default=OUTPUT_FILE)
    ap.add_argument("--no-batch", action="store_true", help="solve the m_p conditions one at a time")
    a = ap.parse_args(argv)
    store = sweep.open_store(a.output)
    try:
        s = sweep.run_sweep(store, RNG_SEED, MAX_TRY, batch_m_p=not a.no_batch)
    except KeyboardInterrupt:
        store.close()
        sweep._log_default("Interrupted - file closed cleanly.")
        raise SystemExit(0)
    store.close()
    sweep._log_default(f"done: {s['computed']} computed, {s['skipped']} skipped, "
                       f"{s['not_converged']} not converged")


if __name__ == "__main__":
    main()
