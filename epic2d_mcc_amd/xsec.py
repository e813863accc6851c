"""Cross-section tables for the null-collision MCC (epic_mcc_setup).

reference_e_ar   the reference's analytic electron-argon cross sections
                 (compute_elastic_CS / compute_excitation_CS /
                 compute_ionization_CS, cross_section.cpp:4-39), tabulated on
                 the log-spaced grid the kernel interpolates.  The
                 reference's enable flags (collision.cpp:28-32 via
                 cross_section.cpp:6,15,27) become the desc's enable[].
synthetic_ar     the MCC workload's synthetic tables (200 log-spaced points,
                 0.01-1000 eV): e-Ar elastic peaking near 1e-19 m^2,
                 excitation above 11.5 eV, ionisation above 15.8 eV, and
                 Ar+-Ar elastic and charge exchange of ~5e-19 m^2 falling as
                 E^-1/2.  No reference counterpart.
"""
import ctypes

import numpy as np

from . import _lib

E_EXC_TH = 11.5  # eV, include/cross_sections.h:14
E_ION_TH = 15.8  # eV, include/cross_sections.h:15
M_AR = 6.63352090e-26


def sigma_elastic(e):
    e = np.asarray(e, dtype=np.float64)
    return 1e-20 * np.abs(6.0 / (1.0 + (e / 0.1) + (e / 0.6) ** 2.0) ** 3.3
                          - 1.1 * e ** 1.4 / (1.0 + (e / 15.0) ** 1.2) / np.sqrt(1.0 + (e / 5.5) ** 2.5
                                                                                 + (e / 60.0) ** 4.1)
                          + 0.05 / (1.0 + e / 10.0) ** 2.0 + 0.01 * e ** 3.0 / (1.0 + (e / 12.0) ** 6.0))


def sigma_excitation(e):
    e = np.asarray(e, dtype=np.float64)
    d = np.maximum(e - E_EXC_TH, 0.0)
    s = 1e-20 * (0.034 * d ** 1.1 * (1.0 + (e / 15.0) ** 2.8) / (1.0 + (e / 23.0) ** 5.5)
                 + 0.023 * d / (1.0 + e / 80.0) ** 1.9)
    return np.where(e > E_EXC_TH, s, 0.0)


def sigma_ionization(e):
    e = np.asarray(e, dtype=np.float64)
    d = np.maximum(e - E_ION_TH, 0.0)
    s = 1e-20 * (970.0 * d / (70.0 + e) ** 2.0 + 0.06 * d ** 2.0 * np.exp(-e / 9))
    return np.where(e > E_ION_TH, s, 0.0)


def log_grid(e_min, e_max, n):
    return np.exp(np.linspace(np.log(e_min), np.log(e_max), n))


def reference_e_ar(n=4096, e_min=0.005, e_max=25000.0):
    """(e_min, e_max, table[5, n]) of the reference's e-Ar formulas."""
    e = log_grid(e_min, e_max, n)
    tab = np.zeros((_lib.MCC_NPROC, n))
    tab[_lib.MCC_E_ELASTIC] = sigma_elastic(e)
    tab[_lib.MCC_E_EXCITATION] = sigma_excitation(e)
    tab[_lib.MCC_E_IONIZATION] = sigma_ionization(e)
    return e_min, e_max, tab


def synthetic_ar(n=200, e_min=0.01, e_max=1000.0):
    e = log_grid(e_min, e_max, n)
    tab = np.zeros((_lib.MCC_NPROC, n))
    tab[_lib.MCC_E_ELASTIC] = 1e-19 * 2.0 * (e / 10.0) / (1.0 + (e / 10.0) ** 2) + 5e-21
    d = np.maximum(e - E_EXC_TH, 0.0)
    tab[_lib.MCC_E_EXCITATION] = np.where(e > E_EXC_TH, 2e-20 * (d / 20.0) / (1.0 + (d / 20.0) ** 1.5), 0.0)
    d = np.maximum(e - E_ION_TH, 0.0)
    tab[_lib.MCC_E_IONIZATION] = np.where(e > E_ION_TH, 3e-20 * (d / 30.0) / (1.0 + (d / 30.0) ** 1.2), 0.0)
    tab[_lib.MCC_I_ELASTIC] = 0.5 * 5e-19 / np.sqrt(e)
    tab[_lib.MCC_I_CX] = 0.5 * 5e-19 / np.sqrt(e)
    return e_min, e_max, tab


def mcc_desc(table, electron_sp=0, ion_sp=1, gas_density=1e21, gas_temp_K=300.0, gas_mass=M_AR,
             enable=(1, 1, 1, 1, 1), ref_selection=1, event_cap=0):
    """epic_mcc_desc for a table from reference_e_ar() / synthetic_ar().  The
    returned desc keeps a reference to its table array."""
    e_min, e_max, tab = table
    arr = np.ascontiguousarray(tab, dtype=np.float64)
    d = _lib.MccDesc(electron_sp=electron_sp, ion_sp=ion_sp, gas_density=gas_density, gas_temp_K=gas_temp_K,
                     gas_mass=gas_mass, e_exc_th=E_EXC_TH, e_ion_th=E_ION_TH, ntab=arr.shape[1], e_min=e_min,
                     e_max=e_max, sigma=arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                     enable=(ctypes.c_int32 * 5)(*enable), ref_selection=ref_selection, event_cap=event_cap)
    d._keep = arr
    return d
