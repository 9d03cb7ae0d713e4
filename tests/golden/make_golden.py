"""Write the golden fixtures under tests/golden/ from the reference's data files.

Run in a container where the reference tree is mounted (read-only) at
/root/reference.  It reads only data: the HDF5 results file of the
reference's parameter sweep (``dda_results/pcas_ocbs_simulated_data.hdf5``,
written by ``run_dda.py:131-141``) and the recorded outputs of
``test_dda.ipynb`` (transcribed below with their cell numbers).  It never
imports or executes reference code.

Outputs:
  pcas_ocbs_simulated_data.npz   every dataset of the HDF5 file, keyed by path
  test_dda_notebook.npz          recorded outputs of test_dda.ipynb
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import hdf5_min  # noqa: E402

REF = os.environ.get("BDDA_REFERENCE", "/root/reference")


def main():
    data = hdf5_min.read_all(os.path.join(REF, "dda_results", "pcas_ocbs_simulated_data.hdf5"))
    np.savez_compressed(os.path.join(HERE, "pcas_ocbs_simulated_data.npz"),
                        **{k.strip("/").replace("/", "__"): v for k, v in data.items()})
    # test_dda.ipynb recorded outputs (cells 4, 5, 6, 8, 9, 10, 11, 14)
    nb = dict(
        r_v_base=0.2, bc_ratio=1.0, ab_ratio=1.0, beta=0.0, shape_seed=1234,
        wl_0=0.834, m_m=1.329,
        m_p_xyz=np.array([1.5 + 0.01j] * 3, dtype=np.complex64),
        lattice_n=np.array([17, 17, 17], dtype=np.int32),
        euler_deg=np.array([
            [81.84096809, 44.68422863, 29.37404504],
            [114.0330023, 170.79860733, 57.56241639],
            [287.05156464, 120.10274156, 122.43606658],
            [243.45168147, 17.26162841, 167.46953533],
            [140.79943822, 79.53113991, 95.91157018],
            [119.81301403, 159.56638548, 293.67950523],
            [215.39115129, 125.54162998, 69.58598014],
            [67.22430682, 58.76511553, 46.60886742],
            [242.19217585, 132.1070694, 32.99931056],
            [339.0490315, 39.624292, 215.48448492]]),
        dpl=20.84953710436821,
        err_printed=np.array([0.0156, 0.0002]), iter_fin=1,
        mie_q=np.array([0.09837689353262594, 0.04493662949384336, 0.1433135230264693]),
        mie_s_fw=np.complex64(0.073400885 + 0.014740863j),
        mie_s_bk=np.complex64(0.00028208116 + 0.0034718874j),
        qabs_relerr=np.array([0.00234075, 0.00082194, 0.00247296, 0.00015358, 0.00059328,
                              0.00013724, 0.00211581, 0.00278785, 0.00257769, 0.00216351]),
        qext_relerr=np.array([0.0016806, 0.00110054, 0.00178967, 0.00048768, 0.00071919,
                              0.00056919, 0.00161458, 0.00207042, 0.00190878, 0.00114165]),
        s_bk=np.array([0.00135109 + 0.00372118j, -0.00141074 + 0.0032077j,
                       0.00144375 + 0.00374011j, -0.00083594 + 0.00331428j,
                       -0.00124069 + 0.00324015j, -0.00053676 + 0.00336732j,
                       0.00109785 + 0.00367712j, 0.00171865 + 0.00379029j,
                       0.00151792 + 0.00375321j, 0.00119338 + 0.00369024j]),
    )
    np.savez_compressed(os.path.join(HERE, "test_dda_notebook.npz"), **nb)


if __name__ == "__main__":
    main()
