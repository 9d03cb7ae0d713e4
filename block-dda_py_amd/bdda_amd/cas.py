"""CAS-v2 post-processing of the forward amplitudes: the PCAS
(polarized complex amplitude sensor) depolarization signal-to-noise ratios
and the non-sphere fraction of a set of orientations.

Follows the reference's ``plot_dda_results.ipynb`` (cells 3-4 and 8): per
orientation, with s = S(0)_theta and p = S(0)_phi,

    ReS_mean  = (Re s + Re p) / 2,        ImS_mean  = (Im s + Im p) / 2,
    ReS_depol = (Re s - Re p) / (Re s + Re p), likewise for Im,
    ReS_depol_rms = sqrt((1 + ReS_depol^2) / 2) * ReS_rms / |ReS_mean|,
    ReS_depol_SNR = ReS_depol / ReS_depol_rms, likewise for Im,
    Tot_depol_SNR2 = ReS_depol_SNR^2 + ImS_depol_SNR^2,

with the instrument's RMS noise ReS_rms = 0.0043 um, ImS_rms = 0.0030 um,
and an orientation counts as non-spherical when Tot_depol_SNR2 exceeds
13.82 (the chi-square 2-dof 0.999 quantile).  The inputs are L values per
condition, so this runs on the host (NumPy) on what the GPU solve returned.
"""
import numpy as np

RES_RMS = 0.0043  # um
IMS_RMS = 0.0030  # um
NONSPHERE_THRESHOLD = 13.82


def pcas_depolarization(S_fw_theta, S_fw_phi, res_rms=RES_RMS, ims_rms=IMS_RMS):
    """Per-orientation PCAS mean amplitude and depolarization SNRs
    (plot_dda_results.ipynb cell 4)."""
    s = np.asarray(S_fw_theta)
    p = np.asarray(S_fw_phi)
    re_mean = (s.real + p.real) / 2
    im_mean = (s.imag + p.imag) / 2
    re_dep = (s.real - p.real) / (s.real + p.real)
    im_dep = (s.imag - p.imag) / (s.imag + p.imag)
    re_snr = re_dep / (np.sqrt((1 + re_dep ** 2) / 2) * res_rms / np.abs(re_mean))
    im_snr = im_dep / (np.sqrt((1 + im_dep ** 2) / 2) * ims_rms / np.abs(im_mean))
    return {"ReS_mean": re_mean, "ImS_mean": im_mean, "ReS_depol": re_dep, "ImS_depol": im_dep,
            "ReS_depol_SNR": re_snr, "ImS_depol_SNR": im_snr, "Tot_depol_SNR2": re_snr ** 2 + im_snr ** 2}


def non_sphere_fraction(tot_depol_snr2, threshold=NONSPHERE_THRESHOLD):
    """Fraction of orientations whose Tot_depol_SNR2 exceeds the threshold
    (plot_dda_results.ipynb cell 8)."""
    v = np.asarray(tot_depol_snr2)
    return np.count_nonzero(v > threshold) / v.size


def cas_v2(store, i_pair=0, i_mp=0, i_bc=0, i_ab=0, i_bt=0, rv_indices=None):
    """The notebook's per-size summary for one sweep condition of a results
    file opened with bdda_amd.sweep.open_store (or an h5py file): a list of
    dicts with r_ve, the Mie S(0), the depolarization quantities and the
    non-sphere fraction, one per r_v index."""
    sd = "target/simulated_data/"
    n_rv = len(np.asarray(store["target/r_v_base_list"][:]))
    out = []
    for i_rv in (range(n_rv) if rv_indices is None else rv_indices):
        idx = (i_pair, i_mp, i_rv, i_bc, i_ab, i_bt)
        d = pcas_depolarization(np.asarray(store[sd + "S_fw_PCAS_theta"][idx + (slice(None),)]),
                                np.asarray(store[sd + "S_fw_PCAS_phi"][idx + (slice(None),)]))
        d.update(ind_rv=i_rv, r_ve=float(np.asarray(store[sd + "r_ve"][(i_rv, i_bc, i_ab, i_bt)])),
                 S_fw_mie=complex(np.asarray(store[sd + "S_fw_PCAS_mie"][idx])),
                 Euler_angles=np.asarray(store[sd + "Euler_angles"][idx]),
                 non_sphere_fraction=non_sphere_fraction(d["Tot_depol_SNR2"]))
        out.append(d)
    return out
