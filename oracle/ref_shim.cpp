// TEST INFRASTRUCTURE ONLY: gives the reference's C++ Pfaffian an unmangled
// name so ctypes can call it. Compiled together with the reference's own
// src/pfaffian_cpu.cpp, where it lies, into oracle/_ref/ (never committed).
// The declaration matches include/pfaffian_cpu.h:5 of the reference.
double pfaffian(double* A, int n);

extern "C" double pfref_pfaffian(double* A, int n) { return pfaffian(A, n); }
