// impurity.hpp -- the impurity-helium pair potentials V(r) (r in Angstrom,
// V in K) that fill the impurity table of potenimpini (potenimpini.f90:11-25),
// and the He-surface potential of the lsurf option.
// Host code only (impurity.cpp, built by the host C++ compiler: several fits
// are summed in binary128 as the reference does).
#pragma once

namespace hetd {
namespace imp {

// Select_Pot (V_impur.f90:84-929): umax for r < r_cutoff, else V(r) of the
// named potential ('null' gives 0 everywhere).  selec is compared after
// trailing blanks are dropped (Fortran Trim).  False for a name this
// library does not know (the reference stops there).
bool select_pot(const char* selec, double r, double r_cutoff, double umax, double* v);

// Names known to select_pot, as one blank-separated string.
const char* known_potentials();

// He-TiO2 surface Morse potentials, K: Morse_HeTiO2_1D(z) (morse.f90:1-16)
// and Morse_HeTiO2_3D(x, y, z) (morse.f90:36-99)
double morse_tio2_1d(double z);
double morse_tio2_3d(double x, double y, double z);

}  // namespace imp
}  // namespace hetd
