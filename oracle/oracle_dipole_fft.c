/*
 * oracle_dipole_fft.c -- FFT dipole field restatement (test oracle only).
 * Filled in with the dipole-fft path (hamiltonian/dipole_fft.cc).
 */
#include "oracle.h"
